/*
 * gcge.h -- umbrella header, as the reference's src/gcge.h.
 */
#ifndef GCGE_H_
#define GCGE_H_

#include "gcge_config.h"
#include "gcge_ops.h"
#include "gcge_para.h"
#include "gcge_solver.h"

#endif
