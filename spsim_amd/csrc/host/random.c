/*
 * Host-side random state of the drop-in library.
 *
 * sp_srand(seed) keys everything random in the library (src/spsim.c:385 seeds
 * libspimage's generator with random_seed):
 *  - p_drand48(): uniform [0,1) stream for molecule builders and random
 *    orientations, a SplitMix64 sequence (Steele, Lea and Flood, "Fast
 *    splittable pseudorandom number generators", OOPSLA 2014);
 *  - the Philox key and per-call pattern counters of the GPU noise stages, so
 *    each generate_poisson_noise / calculate_real_detector_output call draws
 *    fresh, reproducible noise.
 * The reference's Poisson stage used GSL's default seed regardless of
 * random_seed (src/noise.c:14-20); here random_seed controls it too.
 */
#include <stdint.h>

#include "spsim_internal.h"

static uint64_t g_seed = 0;
static uint64_t g_sm_state = 0x853c49e6748fea9bULL;
static uint64_t g_poisson_calls = 0;
static uint64_t g_readout_calls = 0;

void sp_srand(int seed) {
  g_seed = (uint64_t)(uint32_t)seed;
  g_sm_state = 0x9E3779B97F4A7C15ULL ^ (g_seed * 0xD1B54A32D192ED03ULL);
  g_poisson_calls = 0;
  g_readout_calls = 0;
}

static uint64_t splitmix64(void) {
  uint64_t z = (g_sm_state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double p_drand48(void) { return (double)(splitmix64() >> 11) * (1.0 / 9007199254740992.0); }

uint64_t spsim_noise_seed(void) { return g_seed; }

uint64_t spsim_next_poisson_pattern(void) { return g_poisson_calls++; }

uint64_t spsim_next_readout_pattern(void) { return g_readout_calls++; }

uint64_t spsim_next_shot_pattern(void) {
  uint64_t p = g_poisson_calls > g_readout_calls ? g_poisson_calls : g_readout_calls;
  g_poisson_calls = p + 1;
  g_readout_calls = p + 1;
  return p;
}
