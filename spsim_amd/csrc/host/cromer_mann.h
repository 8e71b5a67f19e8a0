/* Cromer-Mann form-factor table (definition and provenance in cromer_mann.c). */
#ifndef SPSIM_CROMER_MANN_H
#define SPSIM_CROMER_MANN_H 1

#define SPSIM_CM_ROWS 100
#define SPSIM_CM_COLS 9

#ifdef __cplusplus
extern "C" {
#endif

extern const float spsim_cromer_mann[SPSIM_CM_ROWS][SPSIM_CM_COLS];

#ifdef __cplusplus
}
#endif

#endif
