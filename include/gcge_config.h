/*
 * gcge_config.h -- scalar types of the GCGE C ABI.
 *
 * Same names and types as the reference's src/gcge_config.h:23-24
 * (GCGE_DOUBLE = double, GCGE_INT = int) so that code written against the
 * reference compiles unchanged.  This build has no MPI: one process drives
 * one GPU.
 */
#ifndef GCGE_CONFIG_H_
#define GCGE_CONFIG_H_

#include <float.h>

#define GCGE_DOUBLE double
#define GCGE_INT int

#endif
