/*
 * random.h - process-global pseudo random numbers used for particle
 * momenta at injection (reference em2d/random.h:21,28).
 *
 * set_rand_seed(w, z): the first argument seeds the "w" MWC half and the
 * second the "z" half, as em2d/random.c:25-29 assigns them (the reference
 * header names the parameters the other way round).
 */
#ifndef RANDOM_H_DROPIN
#define RANDOM_H_DROPIN

#include <stdint.h>

void set_rand_seed(uint32_t m_w_, uint32_t m_z_);
double rand_norm(void);

#endif
