// Adjoint (reverse-mode) passes of the DTS solve: launchers over plain
// device pointers, used by the handle code in dts.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dtsamd {

struct AdjArgs {
  const float* gin;      // cotangent of the pass output
  const void* bnd;       // packed bounds of the pass direction
  int bnarrow;           // bnd has 2-byte elements
  const int2* ab;        // per line: (a, b) collapse limits, see adj_collapse
  const float* state;    // horizontal: pass input state (before the overwrite)
  const float* t;        // target, confidence (horizontal)
  const float* c;
  float* vout;           // cotangent of the pass input
  float* tbar;           // horizontal: accumulated d/dt, d/dc
  float* cbar;
  int W, H;
  int elo, ehi;          // analytic window extents (E + 1, E), clamped to the line
  float lambda, step, theta;
  int ow;                // overwrite active (0 < theta < 1)
  int loss;              // DTS_LOSS_L2 or DTS_LOSS_CHARBONNIER
  double* wS;            // wide windows: line prefixes (a plane of doubles) and
  double* wtot;          // block totals (wide.h); null when not needed
  const int* yext;       // vertical: (lower, upper) window extent per 32-row band and
  int ext_groups;        // 32-column group from the bounds pass; null: the caps
  int acc_first;         // horizontal: first pass, d/dt and d/dc start at zero (not read)
  int acc_last;          // horizontal: last pass, its output (the cotangent of v_0, the
                         // target) is added to d/dt instead of stored
};

// Per line of a line-contiguous int64 ct plane ([nlines][L]): a = first
// position x with ct(x) - ct(0) > sfx (L if none), b = last position x with
// ct(L-1) - ct(x) > sfx (-1 if none).  Positions x < a are the ones whose
// lower window bound collapses, positions x > b the ones whose upper bound
// collapses.  E: search radius floor(sigma) (d >= 1 puts a <= E + 1).
int adj_collapse(hipStream_t s, const int64_t* ct, int L, int nlines, int64_t sfx, int E,
                 int2* ab);
// Both directions in one launch: rows of ctH ([H][W], extent Eh) into abH,
// columns of ctV ([W][H], extent Ev) into abV.
int adj_collapse2(hipStream_t s, const int64_t* ctH, const int64_t* ctV, int W, int H,
                  int64_t sfx, int Eh, int Ev, int2* abH, int2* abV);
// Adjoint of one horizontal data pass (with its overwrites).
int adj_xpass(hipStream_t s, const AdjArgs& a);
// Adjoint of one vertical smoothing pass.
int adj_ypass(hipStream_t s, const AdjArgs& a);
// The horizontal adjoint of one iteration (ax, input ax.gin) and the
// vertical adjoint of the iteration before it (ay, output ay.vout) as one
// launch: the horizontal pass's output stays on chip.  ay.gin is not read;
// ay.vout must differ from ax.gin.  adj_xy_fits: whether the windows allow
// it (else adj_xypass returns DTS_ERR_UNSUPPORTED).
bool adj_xy_fits(const AdjArgs& ax, const AdjArgs& ay);
int adj_xypass(hipStream_t s, const AdjArgs& ax, const AdjArgs& ay);
// dst += src over n floats.
int adj_accumulate(hipStream_t s, float* dst, const float* src, size_t n);

}  // namespace dtsamd
