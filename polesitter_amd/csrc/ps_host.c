/* ps_host.c -- instantiates the C host layer of include/polesitter.h
 * (ps_init / ps_prepare_particles / ps_calc_forces / ps_destroy) into
 * libpolesitter.so, with the threaded build's traversal semantics, the way
 * the reference's own tests and benchmark compile it (tests/main.c:7-9). */
#define PS_MULTITHREADING
#define POLESITTER_IMPLEMENTATION
#include "polesitter.h"
