/*
 * spsim-mi355x drop-in: detector readout (reference include/amplification.h:23-26).
 */
#ifndef _AMPLIFICATION_H_
#define _AMPLIFICATION_H_ 1

#include "config.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/amplification.c:77-114 */
void calculate_real_detector_output(CCD *det, Experiment *experiment);
/* src/amplification.c:54-70 */
void calculate_electrons_per_pixel(CCD *det, Experiment *experiment);
/* src/amplification.c:119-153 */
void calculate_noiseless_detector_output(CCD *det, Experiment *experiment);

#ifdef __cplusplus
}
#endif

#endif
