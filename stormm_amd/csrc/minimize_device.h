// Line-search state machine of the minimizer, shared by the per-step line-advance kernel and
// the persistent minimizer.  All arithmetic is double precision, as in the reference
// (src/MolecularMechanics/hpc_minimization.cu:59-61).
#pragma once

#include "internal.h"

namespace stormm_amd {

struct LineState {
  double l_move, s_move, mfac_a, mfac_b, mfac_c, nrg_a, nrg_b, nrg_c, nrg_d;
};

// Cubic through (0, E0), (m1, E1), (m2, E2), (m3, E3) by Gauss-Jordan elimination on the
// augmented [A | I] matrix, then the choice of the final move,
// src/MolecularMechanics/line_movement.cui:129-338.  Returns the move factor and writes the
// multiplier for the baseline move length.  Every loop is unrolled and the pivot row swap is a
// select on each candidate row, so the matrix is never indexed by a run-time value and stays in
// registers (a run-time index would put it in private memory).
__host__ __device__ inline double cubic_line_move(const double evec[4], const double mvec[4],
                                                  double *new_baseline) {
  double col[8][4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const double xv = mvec[r];
    const double x2 = xv * xv;
    col[0][r] = x2 * xv;
    col[1][r] = x2;
    col[2][r] = xv;
    col[3][r] = 1.0;
#pragma unroll
    for (int c = 4; c < 8; c++) {
      col[c][r] = (c - 4 == r) ? 1.0 : 0.0;
    }
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double maxval = fabs(col[i][i]);
    int maxpos = i;
#pragma unroll
    for (int j = i + 1; j < 4; j++) {
      const double a = fabs(col[i][j]);
      if (a > maxval) {
        maxval = a;
        maxpos = j;
      }
    }
#pragma unroll
    for (int j = i + 1; j < 4; j++) {
      const bool swap = (maxpos == j);
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const double ri = col[c][i];
        const double rj = col[c][j];
        col[c][i] = swap ? rj : ri;
        col[c][j] = swap ? ri : rj;
      }
    }
#pragma unroll
    for (int j = i + 1; j < 4; j++) {
      const double mult = col[i][j] / col[i][i];
#pragma unroll
      for (int c = i; c < 8; c++) {
        col[c][j] -= mult * col[c][i];
      }
    }
  }
#pragma unroll
  for (int i = 3; i >= 0; i--) {
    const double div = col[i][i];
#pragma unroll
    for (int c = 0; c < 8; c++) {
      col[c][i] /= div;
    }
#pragma unroll
    for (int j = i - 1; j >= 0; j--) {
      const double mult = col[i][j];
#pragma unroll
      for (int c = i; c < 8; c++) {
        col[c][j] -= mult * col[c][i];
      }
    }
  }
  double abcd[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    abcd[r] = (col[4][r] * evec[0]) + (col[5][r] * evec[1]) + (col[6][r] * evec[2]) +
              (col[7][r] * evec[3]);
  }
  double move_factor, nbm;
  if (evec[0] < evec[1] && evec[0] < evec[2] && evec[0] < evec[3]) {
    move_factor = -mvec[3];
    nbm = 0.8;
  }
  else {
    const double d0 = 3.0 * abcd[0];
    const double d1 = 2.0 * abcd[1];
    const double d2 = abcd[2];
    const double sqrt_arg = (d1 * d1) - (4.0 * d0 * d2);
    if (sqrt_arg < 0.0) {
      if (evec[0] < evec[3]) {
        move_factor = -mvec[3];
        nbm = 0.8;
      }
      else {
        move_factor = 0.0;
        nbm = 1.2;
      }
    }
    else {
      const double ext_i = (-d1 + sqrt(sqrt_arg)) / (2.0 * d0);
      const double ext_ii = (-d1 - sqrt(sqrt_arg)) / (2.0 * d0);
      const double min_pos = ((2.0 * d0 * ext_i) + d1 > 0.0) ? ext_i : ext_ii;
      if (min_pos <= 0.0) {
        if (evec[3] >= evec[0]) {
          move_factor = -mvec[3];
          nbm = 0.8;
        }
        else {
          move_factor = 0.0;
          nbm = 1.2;
        }
      }
      else if (min_pos < mvec[3]) {
        const double epred = ((((abcd[0] * min_pos) + abcd[1]) * min_pos) + abcd[2]) * min_pos +
                             abcd[3];
        if (epred > evec[0]) {
          move_factor = -mvec[3];
          nbm = 0.8;
        }
        else if (epred < evec[3]) {
          move_factor = min_pos - mvec[3];
          nbm = (min_pos > 0.6) ? 1.05 : 0.95;
        }
        else {
          move_factor = 0.0;
          nbm = 1.05;
        }
      }
      else {
        move_factor = 0.0;
        nbm = 1.05;
      }
    }
  }
  *new_baseline = nbm;
  return move_factor;
}

// Record the energy of the current configuration and return the move factor for line move
// `move` (0..3), updating the state as src/MolecularMechanics/line_movement.cui:67-330 does.
// The caller scales the factor by the move length (l_move for move 0, s_move otherwise).
__host__ __device__ inline double line_move_factor(LineState &ls, int move, double cur) {
  double mf = 0.0;
  if (move == 0) {
    ls.nrg_a = cur;
    mf = 1.0;
    ls.mfac_a = mf;
    ls.s_move = ls.l_move;
  }
  else if (move == 1) {
    ls.nrg_b = cur;
    mf = (cur < ls.nrg_a) ? 1.04 : 0.8;
    ls.mfac_b = mf;
  }
  else if (move == 2) {
    ls.nrg_c = cur;
    mf = (cur < ls.nrg_b) ? ls.mfac_b * 1.03 : ls.mfac_b * 0.85;
    ls.mfac_c = mf;
  }
  else {
    ls.nrg_d = cur;
    const double evec[4] = {ls.nrg_a, ls.nrg_b, ls.nrg_c, ls.nrg_d};
    const double mvec[4] = {0.0, ls.mfac_a, ls.mfac_b + ls.mfac_a,
                            ls.mfac_c + ls.mfac_b + ls.mfac_a};
    double nbm;
    mf = cubic_line_move(evec, mvec, &nbm);
    ls.l_move = ls.s_move * nbm;
  }
  return mf;
}

}  // namespace stormm_amd
