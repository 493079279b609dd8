/* chy_output.h -- declaration shared by the hybrid and solver programs (the entry point itself is
 * part of the C ABI in chybrid.h). */
#ifndef CHY_OUTPUT_H
#define CHY_OUTPUT_H
#include "chybrid.h"
#endif
