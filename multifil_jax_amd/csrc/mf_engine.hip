// Host side of the gfx950 half-sarcomere engine: the C ABI of multifil_amd.h,
// the transition-table kernel, kernel-variant choice and launches.
#include "mf_kernel.h"

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>

namespace mf {
const Variant* variants_a(int& n);
const Variant* variants_b(int& n);
const Variant* variants_c(int& n);
const Variant* variants_d(int& n);
const Variant* variants_e(int& n);
const Variant* variants_f(int& n);
const Variant* variants_g(int& n);
const Variant* variants_h(int& n);
}  // namespace mf

namespace {

// Error text of the last failed call on this thread, in a fixed buffer:
// reporting an error allocates nothing.
thread_local char g_err[512];
__attribute__((format(printf, 1, 2))) int fail(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return 1;
}
#define HIPCHK(x)                                                     \
  do {                                                                \
    hipError_t e_ = (x);                                              \
    if (e_ != hipSuccess) return fail("%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

// Launch shape: B instances x (ktm*27 + 2 kxb 2 nb_max) entries.

__global__ __launch_bounds__(256) void tables_kernel(const mf_launch L, int nb_max) {
  const int ktm = table_sets(L, MF_POP_TM), kxb = table_sets(L, MF_POP_XB);
  const int nxb = kxb * 2 * nb_max;
  const int per = ktm * 27 + 2 * nxb;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int b = (int)(gid / per);
  const int m = (int)(gid % per);
  if (b >= L.B) return;
  const mf_instance I = L.instances[b];
  const mf_lattice G = L.lattices[I.lattice];
  const int nb = G.n_bins;
  const Pops P = pops_of(L, b);
  const float* drv = L.drivers + ((int64_t)b * L.t_stride + I.t0 + L.t_base) * 3;
  const float pCa = drv[1];
  float d = drv[2];
  if (I.flags & MF_INST_DYNAMIC) {
    // dynamic spacing: the kinetics of the first step use the carried spacing
    const StateLayout SL = state_layout(G.n_thick, G.n_crowns, G.n_thin, G.n_sites, G.n_xb_per_crown);
    if (!(I.flags & MF_INST_INIT)) d = (float)*(const double*)(L.state + I.state_off + SL.ls);
  }
  float* tab = L.tables + I.table_off;
  if (m == 0) { tab[0] = pCa; tab[1] = d; tab[2] = d; tab[3] = 1.0f; }
  const double dt = L.cfg.dt;
  if (m < ktm * 27) {
    tm_table_entry(P, m / 27, (double)pCa, dt, m % 27, tab + TAB_TM);
    return;
  }
  int key = m - ktm * 27;
  const bool absorbing = key >= nxb;
  if (absorbing) key -= nxb;
  // key = set * 2 nb_max + k2 over the launch's largest bin count
  const int set = key / (2 * nb_max), k2 = key % (2 * nb_max);
  if (k2 >= 2 * nb) return;
  key = set * 2 * nb + k2;
  const float* centers = L.lat_f32 + G.bin_centers;
  xb_table_entry(P, (double)pCa, (double)d, dt, (double)centers[k2 % nb], key, nb, absorbing,
                 tab + tab_xb(ktm), tab + tab_abs(nb, ktm, kxb));
}

// Fixed-spacing variants first; the dynamic-spacing (bordered solve) ones
// also run fixed-spacing instances, and are picked only for launches that
// carry a dynamic instance.  Global-workspace variants come last.
const Variant* variants(int& n) {
  static Variant all[64];
  static int count = [] {
    int k = 0;
    for (auto fn : {&variants_a, &variants_b, &variants_c, &variants_d, &variants_e, &variants_f,
                    &variants_g, &variants_h}) {
      int m = 0;
      const Variant* g = fn(m);
      for (int i = 0; i < m; ++i) all[k++] = g[i];
    }
    return k;
  }();
  n = count;
  return all;
}

Dims dims_of(const mf_lattice& G, const mf_launch& L) {
  Dims d;
  d.Tk = G.n_thick; d.nc = G.n_crowns; d.Tn = G.n_thin; d.ns = G.n_sites; d.nx = G.n_xb_per_crown;
  d.W = G.width; d.nf = G.n_faces; d.nb = G.n_bins;
  d.nk = d.Tk * d.nc; d.N = d.nk + d.Tn * d.ns; d.X = d.nk * d.nx; d.S = d.Tn * d.ns;
  d.nch = d.Tk + d.Tn;
  d.ktm = table_sets(L, MF_POP_TM);
  d.kxb = table_sets(L, MF_POP_XB);
  return d;
}

// Slots and XB registers a lattice needs at block size bs.
void needs(const Dims& d, int bs, int& ms, int& mx) {
  const int nw = bs / WAVE;
  ms = 0;
  for (int w = 0; w < nw; ++w) {
    int m = 0;
    for (int c = w; c < d.nch; c += nw) m += ((c < d.Tk ? d.nc : d.ns) + WAVE - 1) / WAVE;
    ms = std::max(ms, m);
  }
  mx = (d.X + bs - 1) / bs;
}

// First variant (LDS-resident ones come first) that fits every lattice of
// the launch; *lds its dynamic LDS bytes, *ws its workspace bytes per
// instance (0 unless it is a global-workspace variant).
int pick(const mf_launch* L, const mf_lattice* lats, int n_lats, const Variant** out, int64_t* lds,
         int64_t* ws) {
  int n;
  const Variant* v = variants(n);
  const bool r64 = L->cfg.mode == MF_SOLVER_REFINED;
  const bool dyn = (L->launch_flags & MF_LAUNCH_DYNAMIC) != 0;
  for (int i = 0; i < n; ++i) {
    if (v[i].ref64 != r64 || v[i].dyn != dyn || v[i].bs != L->block) continue;
    bool ok = true;
    int64_t need = 0, wneed = 0;
    for (int g = 0; g < n_lats; ++g) {
      const Dims d = dims_of(lats[g], *L);
      int ms, mx;
      needs(d, v[i].bs, ms, mx);
      // node and head indices travel as 16-bit values
      if (ms > v[i].ms || mx > v[i].mx || d.N > 32767 || d.X > 32767) ok = false;
      // register partner lists hold at most MAXP heads per crown
      if (v[i].ms <= 2 && d.nx > MAXP) ok = false;
      // the bordered solve keeps d f / d d per head in the [N] CG buffer
      if (v[i].dyn && d.X > d.N) ok = false;
      need = std::max(need, v[i].lds(d));
      wneed = std::max(wneed, v[i].ws(d));
    }
    if (ok && need <= 160 * 1024) {
      *out = &v[i];
      *lds = need;
      if (ws) *ws = wneed;
      return 0;
    }
  }
  if (dyn) {
    // the dynamic-spacing variants stop at 24 node slots per lane (8x8)
    for (int g = 0; g < n_lats; ++g) {
      int ms, mx;
      needs(dims_of(lats[g], *L), 1024, ms, mx);
      if (ms > 24)
        return fail("dynamic spacing is not available for this lattice (%d node slots per lane at 1024 "
                    "threads; the dynamic-spacing kernels go up to 24, 8x8): the 32-slot form (10x10) is "
                    "switched off, since its one run stopped responding for a reason not yet found", ms);
    }
  }
  return fail("no kernel variant fits this lattice at block %d (LDS or register slots exceeded)", L->block);
}

// Copy the launch's lattice descriptors into lats[0, cap) and validate the
// launch.  With lattices_host set this touches no device memory.
int fetch_lattices(const mf_launch* L, mf_lattice* lats, int cap) {
  if (L->n_lattices <= 0 || L->B <= 0) return fail("empty launch");
  if (L->n_lattices > cap) return fail("%d lattices in one launch, at most %d", L->n_lattices, cap);
  if (L->lattices_host) {
    std::memcpy(lats, L->lattices_host, sizeof(mf_lattice) * L->n_lattices);
  } else {
    HIPCHK(hipMemcpy(lats, L->lattices, sizeof(mf_lattice) * L->n_lattices, hipMemcpyDeviceToHost));
  }
  for (int g = 0; g < L->n_lattices; ++g) {
    const mf_lattice& G = lats[g];
    if (G.n_bins <= 0 || G.n_bins > MF_MAX_BINS) return fail("n_bins out of range");
    if (G.n_xb_per_crown <= 0 || G.width <= 0) return fail("bad lattice");
  }
  if (L->pop_mode != MF_POP_NONE) {
    if (L->pop_mode != MF_POP_MEAN_FIELD && L->pop_mode != MF_POP_EXPLICIT) return fail("unknown pop_mode");
    if (L->n_pops < 1 || L->n_pops > MF_MAX_POPS) return fail("n_pops out of range");
    if (L->n_pops > 1 && !L->pop_params) return fail("subpopulations need pop_params");
    if (L->pop_mode == MF_POP_MEAN_FIELD && L->n_pops > 1 && !L->pop_fractions)
      return fail("mean-field subpopulations need pop_fractions");
    if (L->pop_mode == MF_POP_EXPLICIT && (L->pop_flags & (MF_POP_XB | MF_POP_TM)) && !L->pop_labels)
      return fail("explicit subpopulations need pop_labels");
  }
  return 0;
}

// Lattices one launch may carry (descriptors are copied to a host array of
// this size, on the stack).
constexpr int MAX_LATTICES = 64;

int check_run_args(const mf_launch* L) {
  if (!L) return fail("null launch");
  if (L->block != 256 && L->block != 512 && L->block != 768 && L->block != 1024)
    return fail("block must be 256, 512, 768 or 1024");
  if (L->cfg.n_cg < 1 || L->cfg.n_cg > 64) return fail("n_cg out of range");
  return 0;
}

// Launch the step kernel v for L, with the per-instance workspace at ws
// (stride ws_stride bytes; unused when the variant keeps everything in LDS).
int launch_run(const mf_launch* L, const mf_lattice* lats, const Variant* v, int64_t lds, uint8_t* ws,
               int64_t ws_stride, void* stream) {
  // transition rows into LDS when they fit beside the instance arrays
  int64_t tab_bytes = 0;
  const int kxb = table_sets(*L, MF_POP_XB);
  for (int g = 0; g < L->n_lattices; ++g)
    tab_bytes = std::max<int64_t>(tab_bytes, 4LL * kxb * (2LL * lats[g].n_bins * 48 + 2LL * lats[g].n_bins * 4));
  mf_launch Lk = *L;
  Lk.workspace = ws;
  Lk.workspace_stride = ws_stride;
  Lk.launch_flags &= ~(LAUNCH_TABLES_IN_LDS | LAUNCH_FACE_POS);
  if (lds + tab_bytes <= 160 * 1024) {
    Lk.launch_flags |= LAUNCH_TABLES_IN_LDS;
    lds += tab_bytes;
  }
  // then the candidate-site positions of the nearest-site search
  int64_t fpos_bytes = 0;
  for (int g = 0; g < L->n_lattices; ++g)
    fpos_bytes = std::max<int64_t>(fpos_bytes, 4LL * lats[g].n_thin * lats[g].n_faces * ((lats[g].width + 3) & ~3));
  fpos_bytes = 16 + ((fpos_bytes + 15) & ~int64_t(15));  // + alignment slack of the kernel's offset
  if (lds + fpos_bytes <= 160 * 1024) {
    Lk.launch_flags |= LAUNCH_FACE_POS;
    lds += fpos_bytes;
  }
  HIPCHK(hipFuncSetAttribute(v->fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  void* args[] = {(void*)&Lk};
  HIPCHK(hipLaunchKernel(v->fn, dim3(L->B), dim3(v->bs), args, (size_t)lds, (hipStream_t)stream));
  HIPCHK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// allocation-free entry point: the workspace of the global-workspace
// variants lives in this file-scope device array.  Every GPU holds its own
// copy until the process ends; a launch uses its first B * stride bytes.
// ---------------------------------------------------------------------------
__device__ __align__(256) uint8_t scratch[MF_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::atomic<uint8_t*> g_scratch[MAX_DEVICES];

// The device address of `scratch` on the current device, looked up on the
// first call made with that device current (the host-side name is only a
// stand-in and must never reach a kernel).
int scratch_ptr(uint8_t** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fail("device %d beyond the %d the library tracks", dev, MAX_DEVICES);
  uint8_t* p = g_scratch[dev].load(std::memory_order_acquire);
  if (!p) {
    void* q = nullptr;
    HIPCHK(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
    if (!q) return fail("hipGetSymbolAddress returned no address for the scratch array");
    p = (uint8_t*)q;
    g_scratch[dev].store(p, std::memory_order_release);
  }
  *out = p;
  return 0;
}

}  // namespace

extern "C" {

int mf_abi_version(void) { return MF_ABI_VERSION; }
const char* mf_last_error(void) { return g_err; }

int64_t mf_state_bytes(int32_t n_thick, int32_t n_crowns, int32_t n_thin, int32_t n_sites,
                       int32_t n_xb_per_crown) {
  return state_layout(n_thick, n_crowns, n_thin, n_sites, n_xb_per_crown).bytes;
}
int64_t mf_table_floats(int32_t n_bins) { return tab_floats(n_bins); }
int64_t mf_table_floats_pop(int32_t n_bins, int32_t n_pop_tm, int32_t n_pop_xb) {
  if (n_pop_tm < 1 || n_pop_xb < 1 || n_pop_tm > MF_MAX_POPS || n_pop_xb > MF_MAX_POPS) return -1;
  return tab_floats(n_bins, n_pop_tm, n_pop_xb);
}

int mf_tables(const mf_launch* L, void* stream) {
  if (!L) return fail("null launch");
  mf_lattice lats[MAX_LATTICES];
  if (int e = fetch_lattices(L, lats, MAX_LATTICES)) return e;
  int nb_max = 0;
  for (int g = 0; g < L->n_lattices; ++g) nb_max = std::max(nb_max, lats[g].n_bins);
  const int64_t total =
      (int64_t)L->B * (table_sets(*L, MF_POP_TM) * 27 + 4 * nb_max * table_sets(*L, MF_POP_XB));
  const int grid = (int)((total + 255) / 256);
  hipLaunchKernelGGL(tables_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, *L, nb_max);
  HIPCHK(hipGetLastError());
  return 0;
}

int mf_run_lds_bytes(const mf_launch* L, int64_t* lds_bytes) {
  if (!L || !lds_bytes) return fail("null argument");
  mf_lattice lats[MAX_LATTICES];
  if (int e = fetch_lattices(L, lats, MAX_LATTICES)) return e;
  const Variant* v;
  return pick(L, lats, L->n_lattices, &v, lds_bytes, nullptr);
}

int mf_workspace_bytes(const mf_launch* L, int64_t* per_instance) {
  if (!L || !per_instance) return fail("null argument");
  mf_lattice lats[MAX_LATTICES];
  if (int e = fetch_lattices(L, lats, MAX_LATTICES)) return e;
  const Variant* v;
  int64_t lds;
  return pick(L, lats, L->n_lattices, &v, &lds, per_instance);
}

int mf_run(const mf_launch* L, void* stream) {
  if (int e = check_run_args(L)) return e;
  mf_lattice lats[MAX_LATTICES];
  if (int e = fetch_lattices(L, lats, MAX_LATTICES)) return e;
  const Variant* v;
  int64_t lds, ws = 0;
  if (int e = pick(L, lats, L->n_lattices, &v, &lds, &ws)) return e;
  if (ws > 0 && (L->workspace == nullptr || L->workspace_stride < ws))
    return fail("this launch needs a device workspace of %lld bytes per instance (mf_workspace_bytes)",
                (long long)ws);
  return launch_run(L, lats, v, lds, L->workspace, L->workspace_stride, stream);
}

int mf_run_noalloc(const mf_launch* L, void* stream) {
  if (int e = check_run_args(L)) return e;
  mf_lattice lats[MAX_LATTICES];
  if (int e = fetch_lattices(L, lats, MAX_LATTICES)) return e;
  const Variant* v;
  int64_t lds, ws = 0;
  if (int e = pick(L, lats, L->n_lattices, &v, &lds, &ws)) return e;
  const int64_t stride = (ws + 255) & ~int64_t(255);
  if (stride > 0 && stride * (int64_t)L->B > MF_NOALLOC_SCRATCH_BYTES)
    return fail("mf_run_noalloc: %d instances need %lld bytes of scratch, more than the %lld of the static array",
                L->B, (long long)(stride * (int64_t)L->B), (long long)MF_NOALLOC_SCRATCH_BYTES);
  uint8_t* p = nullptr;
  if (stride > 0)
    if (int e = scratch_ptr(&p)) return e;
  return launch_run(L, lats, v, lds, p, stride, stream);
}

}  // extern "C"
