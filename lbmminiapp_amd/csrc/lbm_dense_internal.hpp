// lbm_dense_internal.hpp -- the dense engine handle and the host helpers that
// the adjoint module (lbm_adjoint.hip) shares with lbm_dense.hip.
#pragma once
#include "lbm_dense_kernels.hpp"

struct lbm_dense {
  int nx = 0, ny = 0, nz = 0, q = 19, pattern = 1, pull = 0, layout = 0, storage = 0;
  int periodic[3] = {1, 1, 1};
  int staged_mode = 1;  // LDS-staged kernels: 0 never, 1 default cases, 2 every case they cover
  // launch shape of the unstaged time-step kernels (block width, dynamic LDS
  // floor that caps workgroups per CU), and of the SOA AA even kernel, whose
  // accesses are all row-aligned
  int step_bx = 256, even_bx = 256;
  size_t step_lds = 0, even_lds = 0;
  size_t lds_max = 65536;  // the device's LDS limit per workgroup
  bool nt = true;          // streaming (non-temporal) accesses in the time-step kernels
  bool force_terms = false;  // time-step kernels with the Guo terms also at F = 0
  int slab_planes = 0;       // AA steps in pairs as z-slabs of this many planes (0: whole box)
  int odd_seg_w = 0;         // AA odd step as row segments of this width (k_dense_seg_odd; 0: off)
  size_t odd_seg_lds = 0;    // its dynamic LDS floor (caps workgroups per CU)
  int xk = 0;                // interleaved rows: nx / 64 (0: plain rows)
  int odd_ir_bx = 256;       // launch shape of the odd step on interleaved rows
  size_t odd_ir_lds = 0;
  int64_t ncell = 0, nfluid = 0, t = 0;
  size_t esize = 8;
  void* f[2] = {nullptr, nullptr};
  int cur = 0;
  uint32_t* links = nullptr;
  double* omega_field = nullptr;
  double* gamma_field = nullptr;
  double* partial = nullptr;  // mass partial sums + result
  int npartial = 1024;
  hipStream_t stream = nullptr;      // the stream every call enqueues on
  hipStream_t own_stream = nullptr;  // the handle's own stream (stream unless the caller set one)
  int live_graphs = 0;               // graphs captured from this handle and not yet destroyed
  hipEvent_t e0 = nullptr, e1 = nullptr;
  lbm::DenseArgs args{};
  size_t device_bytes = 0;
};

namespace lbm {

// launch shapes of the utility kernels: 256-wide x segments per (y, z) row
dim3 block_for(int nx);
dim3 grid_for(const lbm_dense* h);
// one time-step launch of `kind` from f0 to f1 (f0 == f1 for AA), staged when possible,
// over the z-planes [z0, z0 + nzl) (nzl < 0: the whole box)
lbm_status launch_kind(lbm_dense* h, int kind, void* f0, void* f1, int z0 = 0, int nzl = -1);
// dynamic LDS that leaves room for at most k workgroups per CU (0: no cap)
size_t lds_for_blocks_per_cu(int k);
// device allocation accounted in h->device_bytes
lbm_status dmalloc(lbm_dense* h, void** p, size_t bytes, const char* what);
// lbm_dense_create; push_by_gather = false keeps AB push handles on the
// scatter kernel and the push storage convention, irows_ok = false keeps plain
// rows (the adjoint drives the push kernel on the handle's arrays itself and
// indexes them by cell)
lbm_status dense_create(const lbm_dense_config* cfg, lbm_dense** out, bool push_by_gather,
                        bool irows_ok);

}  // namespace lbm
