// Shared definitions of the HIP layer: the device plan, level tables passed
// to kernels by value, error reporting.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "../dplr_hip.h"

#define PLR_MAXL DPLRH_MAXL

// Per-level table handed to kernels by value (fits the 4 KB argument limit).
struct LevelTab
{
    int L, N, b;
    int R[PLR_MAXL];        // stored rank of level d
    int seglen[PLR_MAXL];   // segment length N >> (d+1)
    int T[PLR_MAXL];        // rows per reduction tile at level d
    int tps[PLR_MAXL];      // reduction tiles per segment
    int task_off[PLR_MAXL + 1];
    int nkg[PLR_MAXL];      // rank groups per level
    int part_off[PLR_MAXL]; // offset of level d in the partial-sum buffer
    int cnt_off[PLR_MAXL + 1]; // offset of level d in the arrival counters
    int w_off[PLR_MAXL + 1];   // offset of level d in the coefficient buffer
    const double *P[PLR_MAXL]; // reduction operand (V for A x, U for A^T x)
    const double *Q[PLR_MAXL]; // expansion operand (U for A x, V for A^T x)
    const double *S[PLR_MAXL];
};

struct dplrh_plan
{
    int N, L, b;
    int R[PLR_MAXL];
    size_t uoff[PLR_MAXL + 1], soff[PLR_MAXL + 1];
    double *U, *V, *S, *D; // device, FP64, layout of dplr_hip.h
    int *rank;             // device copy of plr_flat::rank
    // workspace of the reduce/expand kernels, sized for wk_cols columns
    int wk_cols;
    double *part, *w;
    int *cnt;
    size_t part_per_col, w_per_col, cnt_per_col;
    double *tmp; // op(B) staging for transb
    size_t tmp_cap;
};

void plr_set_error(const char *fmt, ...);

#define PLR_HIP(call)                                                                 \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            plr_set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            return -10;                                                               \
        }                                                                             \
    } while (0)

// Fill the level table for A x (trans == 0) or A^T x (trans == 1).
void plr_level_tab(const dplrh_plan *p, int trans, LevelTab *t);
// Make sure the workspace holds ncols columns.
int plr_workspace(dplrh_plan *p, int ncols);
