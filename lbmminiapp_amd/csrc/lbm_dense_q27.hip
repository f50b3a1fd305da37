// lbm_dense_q27.hip -- instantiates the D3Q27 structured-grid kernels
// (one translation unit per velocity set so they compile in parallel).
#include "lbm_dense_kernels.hpp"

namespace lbm {
DenseLaunchFn pick_dense27(int storage, int layout, int kind, bool mask, bool force, bool nt) {
  return pick_dense_q<27>(storage, layout, kind, mask, force, nt);
}
DenseLaunchFn pick_staged27(int storage, int kind, bool force, int nx, bool mask,
                            size_t lds_max, bool nt) {
  return pick_staged_q<27>(storage, kind, force, nx, mask, lds_max, nt);
}
DenseLaunchFn pick_seg_odd27(int storage, bool force, bool nt) {
  return pick_seg_odd_q<27>(storage, force, nt);
}
}  // namespace lbm
