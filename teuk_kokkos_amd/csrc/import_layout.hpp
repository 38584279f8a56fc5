// Descriptors of the band-validated state import on the GPU
// (pipeline_bands.hpp:108-232).
#pragma once

#include <hip/hip_runtime.h>

#include "device_layout.hpp"

namespace teuk {

// Band operators of one (mode, field) of an imported instance, offsets into
// the import's matrix pool: S (T x n, row-major), A (n x T), and the matrix
// giving the coefficients c with S c = x on range(S) (A, or the Gram-corrected
// B).  n = 0: the field is inactive for the mode (its band is empty).
struct ImportRow {
  long long modal_base;  // evolved rows: offset of the line's block in the modal state, else -1
  int s_off, a_off, c_off, n;
};

struct ImportArgs {
  const ImportRow* rows;   // M * 13, (mode, field)
  const double* mats;
  const cdouble* in;       // the instance's nodal state (mode, field, r, theta)
  cdouble* residual;       // evolved rows: x - S c, same layout
  cdouble* modal;          // the plan's modal state
  double* row_norms;       // per (mode, field, r) row: |x|, then |x - S A x|
  int* any_residual;
  int M, N, T;
};

inline constexpr int kImportThreads = 64;
// widest band an import handles (LDS of the coefficients)
inline constexpr int kMaxImportBand = 256;

// |x| and the off-band residual |x - S A x| of every (mode, field, r) row.
hipError_t launch_band_check(const ImportArgs& a, hipStream_t s);
// Modal coefficients of the evolved rows (fields 0..2 of the modes with a
// line) into the modal state, and their nodal residuals.
hipError_t launch_import_modal(const ImportArgs& a, hipStream_t s);
int import_lds_bytes(int ntheta, int max_band);

}  // namespace teuk
