// pf_host.h -- host-side helpers shared by the library's host files: error and
// allocation helpers, launch geometry, per-kernel event timing and the network handle.
//
// One of the library's host files: pf_api.hip includes them in order into a single
// translation unit together with the kernel files, so every launch is a direct call.

#pragma once

// Learned-preconditioner handle: FP32 master weights with Adam moments, their BF16
// layer matrices, and activation buffers grown to the largest tile count seen.
struct pf_lpc {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  float *w = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
  LpcW bf{};
  bool prepped = false;
  int adam_t = 0;
  long long cap = 0, cap_train = 0;
  __bf16 *X = nullptr, *h1 = nullptr, *h2 = nullptr, *h3 = nullptr, *d3 = nullptr, *d2 = nullptr, *d1 = nullptr;
  float *o = nullptr, *go = nullptr, *scale = nullptr;
  float* part = nullptr;
  long long part_cap = 0;
  float *zbuf = nullptr, *qbuf = nullptr, *gz = nullptr;
  // energy objective: the error vector e and r = A e
  float *ebuf = nullptr, *rbuf = nullptr;
  int objective = PF_LPC_OBJECTIVE_RESIDUAL;
  int power_steps = 0;
  // level-1 training vectors: restricted residual, coarse solution, dL/dg, B S dL/dg, dL/ds
  float* l1[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  long long l1_cap = 0;
  long long cell_cap = 0;
  float* stage = nullptr;
  long long stage_cap = 0;
};

#define HIP_TRY(c, expr)                                                         \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) {                                                      \
      (c)->err = std::string(#expr) + ": " + hipGetErrorString(_e);              \
      return PF_ERR_HIP;                                                         \
    }                                                                            \
  } while (0)

#define LAUNCH_CHECK(c) HIP_TRY(c, hipGetLastError())

namespace {

int fail(pf_ctx* c, int code, const std::string& m) {
  if (c) c->err = m;
  return code;
}

// per-kernel event bracket (active only while pf_kernel_timing is enabled)
long long kbegin(pf_ctx* c, int kid) {
  if (!((c->ktime >> kid) & 1u)) return -1;
  while (c->kev.size() < c->kev_used + 2) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    c->kev.push_back(e);
  }
  size_t i = c->kev_used;
  c->kev_used += 2;
  (void)hipEventRecord(c->kev[i], c->stream);
  c->krec.push_back({kid, i});
  return (long long)i;
}
void kend(pf_ctx* c, long long i) {
  if (i >= 0) (void)hipEventRecord(c->kev[(size_t)i + 1], c->stream);
}
#define TIMED(c, kid, ...)              \
  do {                                  \
    long long _ki = kbegin((c), (kid)); \
    __VA_ARGS__;                        \
    kend((c), _ki);                     \
  } while (0)

// A zeroed device array.  The fill goes to the null stream, which is not ordered with
// the library's non-blocking streams, and may still be running when the call returns:
// it is waited for here, or it could land after the first kernels on a context stream
// have written the array (buffers allocated inside a solve are written at once).
template <typename T>
hipError_t dalloc(T** p, long long n) {
  *p = nullptr;
  if (n <= 0) n = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
  if (e != hipSuccess) return e;
  e = hipMemsetAsync(*p, 0, sizeof(T) * (size_t)n, nullptr);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(nullptr);
}

unsigned grid_for(long long work, long long per_block) {
  long long g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// workgroups for a particle loop split into 8 contiguous XCD shares: a multiple
// of 8, at most cap
unsigned xcd_grid(long long work, long long per_block, unsigned cap) {
  long long g = (work + per_block - 1) / per_block;
  g = ((g + 7) / 8) * 8;
  if (g < 8) g = 8;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// workgroups for the list kernels: a multiple of 8 (one share per XCD), at most
// kMaxPartials (one partial sum each)
unsigned list_grid(long long nlist) {
  long long g = (nlist + 3) / 4;
  g = ((g + 7) / 8) * 8;
  if (g < 8) g = 8;
  if (g > kMaxPartials) g = kMaxPartials;
  return (unsigned)g;
}

const Dims& dims_of(const pf_ctx* c, int field) {
  switch (field) {
    case PF_FIELD_U: case PF_FIELD_MU: case PF_FIELD_SAVED_U: return c->du;
    case PF_FIELD_V: case PF_FIELD_MV: case PF_FIELD_SAVED_V: return c->dv;
    case PF_FIELD_W: case PF_FIELD_MW: case PF_FIELD_SAVED_W: return c->dw;
    default: return c->dc;
  }
}

int field_slot(int field) {
  switch (field) {
    case PF_FIELD_U: return F_U;
    case PF_FIELD_V: return F_V;
    case PF_FIELD_W: return F_W;
    case PF_FIELD_MU: return F_MU;
    case PF_FIELD_MV: return F_MV;
    case PF_FIELD_MW: return F_MW;
    case PF_FIELD_P: return F_P;
    case PF_FIELD_MARKER: return F_MK;
    case PF_FIELD_SAVED_U: return F_SU;
    case PF_FIELD_SAVED_V: return F_SV;
    case PF_FIELD_SAVED_W: return F_SW;
    default: return -1;
  }
}

bool is_pow2(double dx) {
  int e;
  double m = std::frexp(dx, &e);
  return m == 0.5;
}

PhaseK phase_k(const pf_phase* ph) {
  PhaseK k;
  k.rho_l = ph->rho_l;
  k.rho_g = ph->rho_g;
  // etaPhi(pp) * pp.rho_g * pp.rho_tilde_0, phasefield.h:10-11
  k.rmin = std::log(ph->rho_l / ph->rho_g) * ph->rho_g * ph->rho_tilde_0;
  k.den = ph->alpha_phi * ph->rho_tilde_0 * ph->rho_l;
  return k;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

int read_small(pf_ctx* c) {
  HIP_TRY(c, hipMemcpyAsync(c->red_host, c->red, sizeof(double) * 16, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(c->u32_host, c->u32_dev, sizeof(uint32_t) * 16, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return PF_OK;
}

void record(pf_ctx* c, bool timing, int slot) {
  if (timing) hipEventRecord(c->ev[slot], c->stream);
}


// defined in pf_host_lpc.hip; the solver couples to them
int lpc_cw_tiles(pf_ctx* c, const uint32_t* list, long long nlist);
int lpc_cw_weights(pf_ctx* c, bool store);
int lpc_features_alloc(pf_ctx* c);

}  // namespace
