// Device helpers shared by the gfx950 kernels: parameter indices, Philox,
// wave-level scans/reductions, the rate laws and the small-matrix expm.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/multifil_amd.h"

namespace mf {

// DynamicParams row layout (params.py:73-553 field order).
enum Param {
  P_THICK_K = 0, P_THIN_K, P_C_REST_W, P_C_REST_S, P_C_K_W, P_C_K_S, P_G_REST_W, P_G_REST_S,
  P_G_K_W, P_G_K_S, P_TITIN_A, P_TITIN_B, P_TITIN_REST, P_TM_K01, P_TM_K12, P_TM_K23,
  P_TM_K30, P_TM_KEQ01, P_TM_KEQ12, P_TM_KEQ23, P_TM_JC, P_TM_JM, P_R01, P_R12, P_R23,
  P_R34, P_D23, P_D34, P_R40, P_R04, P_R05, P_U_DRX, P_U_LOOSE, P_U_T1, P_U_T2, P_SRX_K0,
  P_SRX_KMAX, P_SRX_B, P_SRX_CA50, P_TEMP, P_SOLVER_TOL, P_PCA, P_ZLINE, P_LS
};
static_assert(P_LS + 1 == MF_N_PARAMS, "parameter row size");

constexpr float RATE_CAP = 10000.0f;
constexpr int N_SQUARINGS_CAP = 18;
constexpr float N_TITIN = 6.0f;
constexpr float XB_REACH = 13.0f;
constexpr int WAVE = 64;

// e^x in double precision for the step kernel's loops: Cody-Waite reduction
// x = n ln2 + r with |r| <= ln2 / 2 (ln 2 split as 6.93147180369123816490e-01
// + 1.90821492927058770002e-10, the fdlibm split, so that n ln2_hi is exact
// for |n| < 2^11), the degree-13 Taylor polynomial of e^r by Horner's rule
// (truncation below 1e-17 relative), then the scaling by 2^n.  Within 2 ulp
// of e^x (tests/test_exp_f64.py checks it against the C library).  On the
// device the coefficients are read through a pointer the compiler cannot see
// through, so in a loop it reloads them (scalar loads) instead of holding
// fourteen doubles in registers across the loop, as it does with the
// library exp's constants (and then spills and reloads them every use).
#define MF_EXP_TAYLOR                                                                             \
  {1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0, 1.0 / 362880.0,     \
   1.0 / 40320.0,      1.0 / 5040.0,      1.0 / 720.0,      1.0 / 120.0,     1.0 / 24.0,         \
   1.0 / 6.0,          0.5,               1.0,              1.0}
static __constant__ double kExpTaylor[14] = MF_EXP_TAYLOR;
__host__ __device__ inline double exp_f64(double x) {
  if (!(x <= 709.782712893384)) return x != x ? x : __builtin_huge_val();
  if (x < -745.1332191019412) return 0.0;
  const double n = rint(x * 1.4426950408889634);
  const double r = fma(-n, 1.90821492927058770002e-10, fma(-n, 6.93147180369123816490e-01, x));
#ifdef __HIP_DEVICE_COMPILE__
  const __attribute__((address_space(4))) double* c = (const __attribute__((address_space(4))) double*)kExpTaylor;
  asm volatile("" : "+s"(c));
#else
  static const double c[14] = MF_EXP_TAYLOR;
#endif
  double p = c[0];
#pragma unroll
  for (int k = 1; k < 14; ++k) p = fma(p, r, c[k]);
  return ldexp(p, (int)n);
}

// atan2(y, x) for y > 0 (a crossbridge's angle: y the lattice spacing, at
// least 1 nm) in double precision.  |x| <= y: pi/2 - atan(x / y); |x| > y:
// atan(y / |x|), mirrored to pi - atan(y / |x|) for x < 0.  Then atan(t),
// |t| <= 1: above tan(pi/8), pi/4 + atan((|t| - 1) / (|t| + 1)), so that the
// odd Taylor series u - u^3/3 + u^5/5 - ... runs on |u| <= tan(pi/8) = 0.4142,
// where 22 terms leave a truncation below 1e-18 (Horner's rule in u^2).
// Within 4 ulp of the C library's atan2 over the spacings and offsets of a
// lattice (tests/test_exp_f64.py).  Coefficients through an opaque pointer on
// the device, as in exp_f64.
#define MF_ATAN_TAYLOR                                                                          \
  {-1.0 / 43.0, 1.0 / 41.0, -1.0 / 39.0, 1.0 / 37.0, -1.0 / 35.0, 1.0 / 33.0, -1.0 / 31.0, 1.0 / 29.0, \
   -1.0 / 27.0, 1.0 / 25.0, -1.0 / 23.0, 1.0 / 21.0, -1.0 / 19.0, 1.0 / 17.0, -1.0 / 15.0, 1.0 / 13.0, \
   -1.0 / 11.0, 1.0 / 9.0,  -1.0 / 7.0,  1.0 / 5.0,  -1.0 / 3.0,  1.0}
static __constant__ double kAtanTaylor[22] = MF_ATAN_TAYLOR;
__host__ __device__ inline double atan2_pos_f64(double y, double x) {
  const double ax = fabs(x);
  const bool near_axis = ax <= y;  // the angle lies in [pi/4, 3pi/4]
  const double t = near_axis ? x / y : y / ax;
  const double at = fabs(t);
  const bool big = at > 0.41421356237309503;
  const double u = big ? (at - 1.0) / (at + 1.0) : at;
#ifdef __HIP_DEVICE_COMPILE__
  const __attribute__((address_space(4))) double* c = (const __attribute__((address_space(4))) double*)kAtanTaylor;
  asm volatile("" : "+s"(c));
#else
  static const double c[22] = MF_ATAN_TAYLOR;
#endif
  const double z = u * u;
  double p = c[0];
#pragma unroll
  for (int k = 1; k < 22; ++k) p = fma(p, z, c[k]);
  double a = fma(u, p, big ? 0.78539816339744831 : 0.0);  // atan(|t|)
  a = t < 0.0 ? -a : a;                                     // atan(t)
  if (near_axis) return 1.5707963267948966 - a;
  return x > 0.0 ? a : 3.1415926535897931 - a;
}

// Per-instance table layout (floats).
constexpr int TAB_HDR = 16;                 // pCa, d_kin, d_abs, valid, ...
constexpr int TAB_TM = TAB_HDR;             // ktm x 27 configs x 4 rows x 4 cumsums
constexpr int TAB_XB = TAB_TM + 27 * 16;    // (ktm = 1) kxb x 2*nb keys x 6 rows x 8 (6 cumsums + pad)
// ktm / kxb: TM / XB table sets, one per population of an explicit
// subpopulation run (labels select the set), else 1.
__host__ __device__ inline int64_t tab_xb(int ktm) { return TAB_TM + (int64_t)ktm * 27 * 16; }
__host__ __device__ inline int64_t tab_abs(int nb, int ktm = 1, int kxb = 1) {
  return tab_xb(ktm) + (int64_t)kxb * 2 * nb * 48;
}
__host__ __device__ inline int64_t tab_floats(int nb, int ktm = 1, int kxb = 1) {
  return tab_abs(nb, ktm, kxb) + (int64_t)kxb * 2 * nb * 4;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Uniform number `idx` of (step, purpose) for an instance: word idx%4 of
// Philox block idx/4, as (bits >> 9) * 2^-23.
__device__ __forceinline__ float philox_uniform(uint64_t seed, int step, int purpose, int idx) {
  const uint4 w = philox4x32_10(make_uint4((uint32_t)(idx >> 2), (uint32_t)step, (uint32_t)purpose, 0u),
                                (uint32_t)seed, (uint32_t)(seed >> 32));
  const int k = idx & 3;
  const uint32_t bits = k == 0 ? w.x : k == 1 ? w.y : k == 2 ? w.z : w.w;
  return (float)(bits >> 9) * 1.1920928955078125e-07f;
}

// philox_uniform(seed, step, 0, idx) and philox_uniform(seed, step, 1, idx)
// for idx = the lane's thread index, at one Philox block per lane instead of
// two: in each quad (idx 4m..4m+3, which share block m) lanes 0-1 compute
// block m of purpose 0 and lanes 2-3 block m of purpose 1, and every lane
// takes its word idx%4 of each by a quad broadcast from lanes 0 and 2.
// All four lanes of every quad must be active.
// (a quad permutation reads a valid lane for every lane, so the move needs
// no "old" operand and mov_dpp spares the copy of one into the destination)
template <int CTRL>
__device__ __forceinline__ uint32_t quad_bcast(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xf, 0xf, false);
}
__device__ __forceinline__ void philox_uniform_pair(uint64_t seed, int step, int idx, float& u0, float& u1) {
  const int r = idx & 3;
  const uint4 w = philox4x32_10(make_uint4((uint32_t)(idx >> 2), (uint32_t)step, r < 2 ? 0u : 1u, 0u),
                                (uint32_t)seed, (uint32_t)(seed >> 32));
  // quad_perm(0,0,0,0) = 0x00 and quad_perm(2,2,2,2) = 0xAA
  const uint32_t a0 = quad_bcast<0x00>(w.x), a1 = quad_bcast<0x00>(w.y), a2 = quad_bcast<0x00>(w.z),
                 a3 = quad_bcast<0x00>(w.w);
  const uint32_t b0 = quad_bcast<0xAA>(w.x), b1 = quad_bcast<0xAA>(w.y), b2 = quad_bcast<0xAA>(w.z),
                 b3 = quad_bcast<0xAA>(w.w);
  const uint32_t x0 = r == 0 ? a0 : r == 1 ? a1 : r == 2 ? a2 : a3;
  const uint32_t x1 = r == 0 ? b0 : r == 1 ? b1 : r == 2 ? b2 : b3;
  u0 = (float)(x0 >> 9) * 1.1920928955078125e-07f;
  u1 = (float)(x1 >> 9) * 1.1920928955078125e-07f;
}

// ---------------------------------------------------------------------------
// wave-level primitives (wave64, DPP)
//
// Inclusive prefix: Hillis-Steele inside each 16-lane row with row_shr:1/2/4/8,
// then row_bcast:15 (rows 1, 3 take lane 15 of the row below) and
// row_bcast:31 (rows 2, 3 take lane 31).  Totals are read from lane 63.
// Each step is one DPP move plus the operation: lanes whose source lies
// outside the row, and rows outside the row mask, keep the "old" operand,
// which is the operation's identity (0 for sums, -inf for maxima), so no
// lane predicate is needed.
// ---------------------------------------------------------------------------
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_shift(float v, float id) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(id), __float_as_int(v), CTRL, ROWS, 0xf, false));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_shift(double v, double id) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(id), __double2loint(v), CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(id), __double2hiint(v), CTRL, ROWS, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float read_lane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double read_lane(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

struct OpAdd {
  template <typename T> __device__ T operator()(T a, T b) const { return a + b; }
  template <typename T> __device__ static T identity() { return T(0); }
};
struct OpMax {
  template <typename T> __device__ T operator()(T a, T b) const { return fmax(a, b); }
  template <typename T> __device__ static T identity() { return -T(INFINITY); }
};

// inclusive prefix over the 64 lanes (lane 0 first)
template <typename T, typename Op = OpAdd>
__device__ __forceinline__ T wave_prefix(T v, int lane, Op op = Op()) {
  (void)lane;
  const T id = Op::template identity<T>();
  v = op(v, dpp_shift<0x111, 0xf>(v, id));   // row_shr:1
  v = op(v, dpp_shift<0x112, 0xf>(v, id));   // row_shr:2
  v = op(v, dpp_shift<0x114, 0xf>(v, id));   // row_shr:4
  v = op(v, dpp_shift<0x118, 0xf>(v, id));   // row_shr:8
  v = op(v, dpp_shift<0x142, 0xa>(v, id));   // row_bcast:15 into rows 1, 3
  v = op(v, dpp_shift<0x143, 0xc>(v, id));   // row_bcast:31 into rows 2, 3
  return v;
}
// exclusive prefix (0 on lane 0), from the inclusive one shifted by a lane
template <typename T>
__device__ __forceinline__ T wave_prefix_excl(T incl, int lane) {
  (void)lane;
  return dpp_shift<0x138, 0xf>(incl, T(0));  // wave_shr:1; lane 0 has no source
}
// Reduction over each 16-lane row (row_shr:1/2/4/8), total in the row's
// lane 15.
template <typename T, typename Op = OpAdd>
__device__ __forceinline__ T row_reduce16(T v, Op op = Op()) {
  const T id = Op::template identity<T>();
  v = op(v, dpp_shift<0x111, 0xf>(v, id));
  v = op(v, dpp_shift<0x112, 0xf>(v, id));
  v = op(v, dpp_shift<0x114, 0xf>(v, id));
  v = op(v, dpp_shift<0x118, 0xf>(v, id));
  return v;
}
// wave-uniform reductions
template <typename T>
__device__ __forceinline__ T wave_sum(T v, int lane) { return read_lane(wave_prefix(v, lane, OpAdd()), 63); }
template <typename T>
__device__ __forceinline__ T wave_max(T v, int lane) { return read_lane(wave_prefix(v, lane, OpMax()), 63); }

// ---------------------------------------------------------------------------
// rate laws (rate_functions.py) and generators (transitions.py)
// ---------------------------------------------------------------------------
// Rate laws and the expm are templated on the arithmetic type: the table
// kernel evaluates them in double (the tables are built once per pCa /
// spacing), so the stored float32 cumulative rows carry only their own
// rounding.
template <typename T>
__device__ __forceinline__ T ca_conc(T pCa) { return pow(T(10), -pCa); }

// Thermal energy in pN nm (transitions.py:708).
template <typename T>
__device__ __forceinline__ T kT_pNnm(T temp_c) {
  return (T(1.3810e-23) * (temp_c + T(273.15))) * T(1e21);
}

// 4x4 TM generator of neighbour configuration cfg = 9 n2 + 3 n3 + n_closed
// (transitions.py:434-513, rate_functions.py:165-311).
// ACC: Q += w * generator (the mean-field blend of populations,
// transitions.py:585-597), else Q = generator.
template <typename T, bool ACC = false>
__device__ inline void tm_generator(const float* prm, T pCa, int cfg, T Q[4][4], T w = T(1)) {
  const T n2 = (T)(cfg / 9), n3 = (T)((cfg / 3) % 3), nc = (T)(cfg % 3);
  const T JC = prm[P_TM_JC], JM = prm[P_TM_JM];
  const T h = JC * n2 + JM * n3 - T(0.5) * (JC + JM) * nc;
  const T fwd = exp(T(0.5) * h), bwd = exp(T(-0.5) * h);
  const T ca = ca_conc<T>(pCa);
  const T k01 = T(prm[P_TM_K01]) * ca * fwd;
  const T k12 = T(prm[P_TM_K12]) * fwd;
  const T k23 = T(prm[P_TM_K23]) * fwd;
  const T k30 = prm[P_TM_K30];
  const T k10 = (T(prm[P_TM_K01]) / T(prm[P_TM_KEQ01])) * bwd;
  const T k21 = (T(prm[P_TM_K12]) / T(prm[P_TM_KEQ12])) * bwd;
  const T k32 = (T(prm[P_TM_K23]) / T(prm[P_TM_KEQ23])) * bwd;
  if (!ACC) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) Q[i][j] = T(0);
  }
  auto put = [&](int i, int j, T v) { if (ACC) Q[i][j] += w * v; else Q[i][j] = v; };
  put(0, 0, -k01); put(0, 1, k01);
  put(1, 0, k10); put(1, 1, -(k10 + k12)); put(1, 2, k12);
  put(2, 1, k21); put(2, 2, -(k21 + k23)); put(2, 3, k23);
  put(3, 0, k30); put(3, 2, k32); put(3, 3, -(k30 + k32));
}

// The eleven XB rates at axial offset x, spacing d, permissiveness perm
// (transitions.py:632-804, rate_functions.py:318-753).
template <typename T>
struct XbRates { T r01, r04, r05, r10, r12, r21, r23, r32, r34, r40, r50; };

// The parts of the XB rates that depend on neither the head's offset nor
// the spacing: kT (and 1/kT), the calcium-dependent SRX rate r50, and the
// constant factors e^(U_loose - U_DRX) and e^(U_tight_2 - U_tight_1) of the
// reverse rates r10 and r32.
template <typename T>
struct XbPre { T kt, ikt, r50, e10, e32; };

template <typename T>
__device__ inline XbPre<T> xb_pre(const float* prm, T pCa) {
  XbPre<T> p;
  p.kt = kT_pNnm<T>(T(prm[P_TEMP]));
  p.ikt = T(1) / p.kt;
  const T ca = ca_conc<T>(pCa), b = prm[P_SRX_B];
  const T cab = pow(ca, b);
  p.r50 = T(prm[P_SRX_K0]) + ((T(prm[P_SRX_KMAX]) - T(prm[P_SRX_K0])) * cab) / (pow(T(prm[P_SRX_CA50]), b) + cab);
  p.e10 = exp(T(prm[P_U_LOOSE]) - T(prm[P_U_DRX]));
  p.e32 = exp(T(prm[P_U_T2]) - T(prm[P_U_T1]));
  return p;
}

// The reference forms the reverse rates on a log scale:
// r10 = exp(log(r01 + 0.005) - (U_DRX - U_loose)) (rate_functions.py:384-387),
// r21 = exp(log(r12) - (U_loose - U_t1)) (:464-466),
// r32 = exp(min(log(r23 + 1e-30) + (U_t2 - U_t1), log(cap))) (:559-560).
// Here the same products are formed directly, with
// e^(U_loose - U_DRX) e^(E_w) = e10 / e^(-E_w) and e^(U_t2 - U_t1) = e32
// (the strong-state energies cancel): five exponentials and no logarithm
// instead of seven and three.  The results agree to rounding (relative
// 1e-15); the caps, the NaN rules and the underflow ends are those of the
// log-scale forms.
template <typename T>
__device__ inline XbRates<T> xb_rates(const float* prm, const XbPre<T>& pre, T d, T x, T perm) {
  const T cap = T(RATE_CAP);
  const T r = sqrt(x * x + d * d);
  T th;
  if constexpr (std::is_same<T, double>::value) th = atan2_pos_f64(d, x); else th = atan2(d, x);
  const T ikt = pre.ikt;
  const T dgw = r - T(prm[P_G_REST_W]), dcw = th - T(prm[P_C_REST_W]);
  const T dgs = r - T(prm[P_G_REST_S]), dcs = th - T(prm[P_C_REST_S]);
  const T Ew = (T(0.5) * T(prm[P_G_K_W]) * (dgw * dgw) + T(0.5) * T(prm[P_C_K_W]) * (dcw * dcw)) * ikt;
  const T Es = (T(0.5) * T(prm[P_G_K_S]) * (dgs * dgs) + T(0.5) * T(prm[P_C_K_S]) * (dcs * dcs)) * ikt;
  const T Ediff = Ew - Es;
  const T Uloose = T(prm[P_U_LOOSE]) + Ew;
  const T Ut1 = T(prm[P_U_T1]) + Es;
  const T fs = T(prm[P_G_K_S]) * dgs + (T(1) / r) * T(prm[P_C_K_S]) * dcs;
  XbRates<T> k;
  const T emw = exp(-Ew);
  T r01 = perm * T(prm[P_R01]) * emw;
  if (isnan(r01)) r01 = T(0);
  // e^(-E_w) = 0 (E_w > 745) makes the factor infinite and r10 = cap, where
  // the log-scale form's exponential overflows
  T r10 = fmin((r01 + T(0.005)) * (pre.e10 / emw), cap);
  if (isnan(r10)) r10 = cap;
  T r12 = T(prm[P_R12]) * exp(fmin(Ediff, T(30)) / T(2));
  if (isnan(r12)) r12 = T(0);
  k.r01 = r01;
  k.r10 = r10;
  k.r12 = r12;
  // r12 = 0 (underflow) has log -inf and r21 = 0 in the log-scale form
  k.r21 = r12 > T(0) ? fmin(r12 * exp(Ut1 - Uloose), cap) : T(0);
  k.r23 = fmin(T(prm[P_R23]) * exp(-fs * T(prm[P_D23]) * ikt), cap);
  k.r32 = fmin((k.r23 + T(1e-30)) * pre.e32, cap);
  k.r34 = fmin(T(prm[P_R34]) * exp(fs * T(prm[P_D34]) * ikt), cap);
  k.r40 = prm[P_R40];
  k.r04 = prm[P_R04];
  k.r05 = prm[P_R05];
  k.r50 = pre.r50;
  return k;
}

template <typename T>
__device__ inline XbRates<T> xb_rates(const float* prm, T pCa, T d, T x, T perm) {
  return xb_rates<T>(prm, xb_pre<T>(prm, pCa), d, x, perm);
}

// Entry (i, j) of the 6x6 XB generator from its rates; the absorbing variant
// has rows 0 and 4 zeroed (transitions.py:1077).  Layout transitions.py:
// 283-298, diagonals :789-794.
template <typename T>
__device__ __forceinline__ T xb_gen_entry(const XbRates<T>& k, int i, int j, bool absorbing) {
  switch (i) {
    case 0: return absorbing ? T(0) : j == 0 ? -(k.r01 + k.r04 + k.r05) : j == 1 ? k.r01 : j == 4 ? k.r04 : j == 5 ? k.r05 : T(0);
    case 1: return j == 0 ? k.r10 : j == 1 ? -(k.r10 + k.r12) : j == 2 ? k.r12 : T(0);
    case 2: return j == 1 ? k.r21 : j == 2 ? -(k.r21 + k.r23) : j == 3 ? k.r23 : T(0);
    case 3: return j == 2 ? k.r32 : j == 3 ? -(k.r32 + k.r34) : j == 4 ? k.r34 : T(0);
    case 4: return absorbing ? T(0) : j == 0 ? k.r40 : j == 4 ? -k.r40 : T(0);
    default: return j == 0 ? k.r50 : j == 5 ? -k.r50 : T(0);
  }
}

// Row r of the XB generator (r may differ between the lanes of a wave: the
// row is picked by selects, not by a branch).
template <typename T>
__device__ __forceinline__ void xb_gen_row(const XbRates<T>& k, int r, bool absorbing, T out[6]) {
  const T z = T(0);
  const T d0 = absorbing ? z : -(k.r01 + k.r04 + k.r05), d1 = -(k.r10 + k.r12), d2 = -(k.r21 + k.r23);
  const T d3 = -(k.r32 + k.r34), d4 = absorbing ? z : -k.r40, d5 = -k.r50;
  const T a01 = absorbing ? z : k.r01, a04 = absorbing ? z : k.r04, a05 = absorbing ? z : k.r05;
  const T a40 = absorbing ? z : k.r40;
  out[0] = r == 0 ? d0 : r == 1 ? k.r10 : r == 4 ? a40 : r == 5 ? k.r50 : z;
  out[1] = r == 0 ? a01 : r == 1 ? d1 : r == 2 ? k.r21 : z;
  out[2] = r == 1 ? k.r12 : r == 2 ? d2 : r == 3 ? k.r32 : z;
  out[3] = r == 2 ? k.r23 : r == 3 ? d3 : z;
  out[4] = r == 0 ? a04 : r == 3 ? k.r34 : r == 4 ? d4 : z;
  out[5] = r == 0 ? a05 : r == 5 ? d5 : z;
}

// 6x6 XB generator (transitions.py:632-804).  ACC: Q += w * generator (the
// mean-field blend, transitions.py:925-975), else Q = generator.
template <typename T, bool ACC = false>
__device__ inline void xb_generator(const float* prm, T pCa, T d, T x, T perm, bool absorbing,
                                    T Q[6][6], T w = T(1)) {
  const XbRates<T> k = xb_rates<T>(prm, pCa, d, x, perm);
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const T v = xb_gen_entry<T>(k, i, j, absorbing);
      if (ACC) Q[i][j] += w * v; else Q[i][j] = v;
    }
}

// ---------------------------------------------------------------------------
// expm by scaling and squaring, the reference's variant (transitions.py:99-190)
// ---------------------------------------------------------------------------
template <int N, typename T>
__device__ __forceinline__ void matmul(const T A[N][N], const T B[N][N], T C[N][N]) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      T acc = A[i][0] * B[0][j];
#pragma unroll
      for (int k = 1; k < N; ++k) acc = fma(A[i][k], B[k][j], acc);
      C[i][j] = acc;
    }
}

// A <- expm(A).  ||A/2^s||_inf <= 1/2 with s from the float32 norm, as the
// reference computes it; Pade-type rational approximant with the reference's
// coefficients; (V-U) X = (V+U) by Gauss-Jordan without pivoting (V-U =
// I - A_s/2 + O(A_s^2) is diagonally dominant for ||A_s|| <= 1/2);
// min(s, 18) squarings; with RENORM (a generator's exponential: a
// stochastic matrix) rows renormalised and NaN -> 1/N.
template <int N, typename T, bool RENORM = true>
__device__ inline void expm_ref(T A[N][N]) {
  T nrm = T(0);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    T rs = T(0);
#pragma unroll
    for (int j = 0; j < N; ++j) rs += fabs(A[i][j]);
    nrm = fmax(nrm, rs);
  }
  int s = (int)ceilf(log2f((float)nrm / 0.5f + 1e-10f));
  s = s < 0 ? 0 : s;
  const T sc = ldexp(T(1), -s);
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) A[i][j] *= sc;
  const T b0 = T(1), b1 = T(1) / T(2), b2 = T(1) / T(9), b3 = T(1) / T(72), b4 = T(1) / T(1008),
          b5 = T(1) / T(30240), b6 = T(1) / T(1209600);
  T A2[N][N], A4[N][N], V[N][N], W[N][N];
  matmul<N, T>(A, A, A2);
  matmul<N, T>(A2, A2, A4);
  matmul<N, T>(A4, A2, V);  // A6
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const T id = (i == j) ? T(1) : T(0);
      V[i][j] = b0 * id + b2 * A2[i][j] + b4 * A4[i][j] + b6 * V[i][j];
      W[i][j] = b1 * id + b3 * A2[i][j] + b5 * A4[i][j];
    }
  matmul<N, T>(W, A, A2);  // U = W A (= A W: W is a polynomial in A)
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      W[i][j] = V[i][j] - A2[i][j];
      A[i][j] = V[i][j] + A2[i][j];
    }
  // Gauss-Jordan with the pivot rows left unscaled until the end: step k
  // removes column k from every other row with the multiplier W[i][k] /
  // W[k][k]; then each row is scaled by the reciprocal of its pivot (as it
  // was at its own step)
  T ipiv[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const T inv = T(1) / W[k][k];
    ipiv[k] = inv;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (i == k) continue;
      const T f = W[i][k] * inv;
#pragma unroll
      for (int j = 0; j < N; ++j) { W[i][j] = fma(-f, W[k][j], W[i][j]); A[i][j] = fma(-f, A[k][j], A[i][j]); }
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) A[i][j] *= ipiv[i];
  const int nsq = s < N_SQUARINGS_CAP ? s : N_SQUARINGS_CAP;
  for (int q = 0; q < nsq; ++q) {
    matmul<N, T>(A, A, A2);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) A[i][j] = A2[i][j];
  }
  if (!RENORM) return;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    T rs = T(0);
#pragma unroll
    for (int j = 0; j < N; ++j) rs += A[i][j];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const T v = A[i][j] / rs;
      A[i][j] = isnan(v) ? T(1) / T(N) : v;
    }
  }
}

// The three entries of exp(Q_abs dt) the metrics use (P[3][4], P[2][0],
// P[3][0]; Q_abs: the XB generator with rows 0 and 4 zeroed, states 0 and 4
// absorbing, transitions.py:1029-1078), from its transient block.  States
// 1-3 leak only to 0 (r10, from 1) and to 4 (r34, from 3) and never reach
// 5, so with A = T dt (T: the rates among states 1-3, as in Q_abs) the
// absorption probabilities within dt are A^-1 (e^A - I) (R dt), R the two
// exit columns: a 3x3 exponential and one 3x3 solve with two right-hand
// sides (partial pivoting) instead of a 6x6 exponential.  The same numbers
// to FP64 rounding; A is nonsingular (r10 >= the 0.005-floored exit).
// a = A (3x3, already times dt), e10 = r10 dt, e34 = r34 dt.
template <typename T>
__device__ inline void absorbing_probs(const T a[3][3], T e10, T e34, T out[3]) {
  T E[3][3], M[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) { E[i][j] = a[i][j]; M[i][j] = a[i][j]; }
  expm_ref<3, T, false>(E);
  // right-hand sides (E - I) e0 e10 and (E - I) e2 e34
  T y[3], z[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    y[i] = (E[i][0] - (i == 0 ? T(1) : T(0))) * e10;
    z[i] = (E[i][2] - (i == 2 ? T(1) : T(0))) * e34;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    // partial pivoting among rows k..2 (selects, no branches)
#pragma unroll
    for (int i = k + 1; i < 3; ++i) {
      const bool sw = fabs(M[i][k]) > fabs(M[k][k]);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const T u = M[k][j], v = M[i][j];
        M[k][j] = sw ? v : u;
        M[i][j] = sw ? u : v;
      }
      const T uy = y[k], vy = y[i], uz = z[k], vz = z[i];
      y[k] = sw ? vy : uy; y[i] = sw ? uy : vy;
      z[k] = sw ? vz : uz; z[i] = sw ? uz : vz;
    }
    const T inv = T(1) / M[k][k];
#pragma unroll
    for (int i = k + 1; i < 3; ++i) {
      const T f = M[i][k] * inv;
#pragma unroll
      for (int j = k; j < 3; ++j) M[i][j] = fma(-f, M[k][j], M[i][j]);
      y[i] = fma(-f, y[k], y[i]);
      z[i] = fma(-f, z[k], z[i]);
    }
  }
#pragma unroll
  for (int k = 2; k >= 0; --k) {
#pragma unroll
    for (int j = k + 1; j < 3; ++j) {
      y[k] = fma(-M[k][j], y[j], y[k]);
      z[k] = fma(-M[k][j], z[j], z[k]);
    }
    y[k] /= M[k][k];
    z[k] /= M[k][k];
  }
  out[0] = z[2];  // P[3][4]
  out[1] = y[1];  // P[2][0]
  out[2] = y[2];  // P[3][0]
}

// The transient block of a rate set, times dt, accumulated with weight w
// (ACC: the mean-field blend, transitions.py:925-975) or assigned.
template <typename T, bool ACC>
__device__ __forceinline__ void absorbing_block(const XbRates<T>& k, T dt, T w, T a[3][3], T& e10, T& e34) {
  const T b[3][3] = {{-(k.r10 + k.r12), k.r12, T(0)},
                     {k.r21, -(k.r21 + k.r23), k.r23},
                     {T(0), k.r32, -(k.r32 + k.r34)}};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) a[i][j] = ACC ? a[i][j] + w * (b[i][j] * dt) : b[i][j] * dt;
  e10 = ACC ? e10 + w * (k.r10 * dt) : k.r10 * dt;
  e34 = ACC ? e34 + w * (k.r34 * dt) : k.r34 * dt;
}

// ---------------------------------------------------------------------------
// The same expm, computed by the four lanes of a quad: lane l owns rows
// 2 min(l, 2) and 2 min(l, 2) + 1 (lane 3 repeats lane 2's rows), and a
// matrix product takes the other lanes' rows by DPP quad broadcasts.  Every
// row sees the operations of expm_ref in the same order (products with k
// ascending, the same Gauss-Jordan steps; the pivot row's own update, a
// multiply-add with multiplier 0, leaves every entry but a -0 unchanged),
// so the result is expm_ref's bit for bit up to the sign of a zero entry.
// All four lanes of the quad must be active.
// ---------------------------------------------------------------------------
template <int Q>
__device__ __forceinline__ double quad_bcast_d(double v) {
  constexpr int C = Q | (Q << 2) | (Q << 4) | (Q << 6);
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), C, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), C, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Entry (k, j) of the XB generator can be nonzero (xb_gen_row): row 0 at
// columns 0, 1, 4, 5; row 1 at 0-2; row 2 at 1-3; row 3 at 2-4; row 4 at 0
// and 4; row 5 at 0 and 5.  A product whose right operand has this pattern
// (SPB: the scaled generator) skips the other entries' broadcasts and
// multiply-adds, which leave the sums unchanged (but for the sign of a zero).
constexpr bool xb_gen_nz(int k, int j) {
  return k == 0 ? (j == 0 || j == 1 || j == 4 || j == 5)
       : k == 1 ? (j <= 2)
       : k == 2 ? (j >= 1 && j <= 3)
       : k == 3 ? (j >= 2 && j <= 4)
       : k == 4 ? (j == 0 || j == 4)
                : (j == 0 || j == 5);
}

// C = A B for the row pair of this lane (A, C: this lane's rows; B: the
// lane's rows of B).  Row k of B comes from quad lane k / 2, one row at a
// time.  C may alias A or B only through the caller's copy.
template <int K, bool SPB>
__device__ __forceinline__ void quad_fma_row(const double A[2][6], const double B[2][6], double acc[2][6]) {
  double b[6];
#pragma unroll
  for (int j = 0; j < 6; ++j)
    if (!SPB || xb_gen_nz(K, j)) b[j] = quad_bcast_d<K / 2>(B[K & 1][j]);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      if (SPB && !xb_gen_nz(K, j)) {
        if (K == 0) acc[i][j] = 0.0;
      } else {
        acc[i][j] = K == 0 ? A[i][0] * b[j] : fma(A[i][K], b[j], acc[i][j]);
      }
    }
  // keep the scheduler from hoisting the next row's broadcasts above these
  // FMAs: all six rows in flight at once would not fit the register file
  __builtin_amdgcn_sched_barrier(0);
}
template <bool SPB = false>
__device__ __forceinline__ void quad_matmul6(const double A[2][6], const double B[2][6], double C[2][6]) {
  double acc[2][6];
  quad_fma_row<0, SPB>(A, B, acc);
  quad_fma_row<1, SPB>(A, B, acc);
  quad_fma_row<2, SPB>(A, B, acc);
  quad_fma_row<3, SPB>(A, B, acc);
  quad_fma_row<4, SPB>(A, B, acc);
  quad_fma_row<5, SPB>(A, B, acc);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) C[i][j] = acc[i][j];
}

// A <- expm(A) for a 6x6 matrix held as row pairs by a quad; l = lane & 3.
__device__ inline void expm_quad6(double A[2][6], int l) {
  const int rp = l < 2 ? l : 2;  // row pair 0, 1 or 2
  // infinity norm: row sums in order, then the maximum over the quad
  double nrm = 0.0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    double rs = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) rs += fabs(A[i][j]);
    nrm = fmax(nrm, rs);
  }
  nrm = fmax(fmax(quad_bcast_d<0>(nrm), quad_bcast_d<1>(nrm)), quad_bcast_d<2>(nrm));
  int s = (int)ceilf(log2f((float)nrm / 0.5f + 1e-10f));
  s = s < 0 ? 0 : s;
  const double sc = ldexp(1.0, -s);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) A[i][j] *= sc;
  const double b0 = 1.0, b1 = 1.0 / 2.0, b2 = 1.0 / 9.0, b3 = 1.0 / 72.0, b4 = 1.0 / 1008.0,
               b5 = 1.0 / 30240.0, b6 = 1.0 / 1209600.0;
  double A2[2][6], A4[2][6], A6[2][6];
  quad_matmul6<true>(A, A, A2);
  quad_matmul6(A2, A2, A4);
  quad_matmul6(A4, A2, A6);
  // V over A6 and W over A4, entry by entry (each read once, before it is
  // overwritten)
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double id = (2 * rp + i == j) ? 1.0 : 0.0;
      const double v = b0 * id + b2 * A2[i][j] + b4 * A4[i][j] + b6 * A6[i][j];
      const double w = b1 * id + b3 * A2[i][j] + b5 * A4[i][j];
      A6[i][j] = v;
      A4[i][j] = w;
    }
  double (&V)[2][6] = A6;
  double (&W)[2][6] = A4;
  quad_matmul6<true>(W, A, A2);  // U = W A (= A W: W is a polynomial in A)
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      W[i][j] = V[i][j] - A2[i][j];
      A[i][j] = V[i][j] + A2[i][j];
    }
  // Gauss-Jordan as expm_ref's, pivot rows unscaled until the end: the
  // pivot row's own lane takes multiplier 0 (a select on the multiplier,
  // not on every entry; fma(-0, p, w) returns w) and keeps the pivot's
  // reciprocal for the final scaling
  double ipiv[2] = {0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double pw[6], pa[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const int kr = k & 1;
      pw[j] = (k >> 1) == 0 ? quad_bcast_d<0>(W[kr][j]) : (k >> 1) == 1 ? quad_bcast_d<1>(W[kr][j])
                                                                        : quad_bcast_d<2>(W[kr][j]);
      pa[j] = (k >> 1) == 0 ? quad_bcast_d<0>(A[kr][j]) : (k >> 1) == 1 ? quad_bcast_d<1>(A[kr][j])
                                                                        : quad_bcast_d<2>(A[kr][j]);
    }
    const double inv = 1.0 / pw[k];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const bool own = 2 * rp + i == k;
      ipiv[i] = own ? inv : ipiv[i];
      const double f = own ? 0.0 : W[i][k] * inv;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        W[i][j] = fma(-f, pw[j], W[i][j]);
        A[i][j] = fma(-f, pa[j], A[i][j]);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) A[i][j] *= ipiv[i];
  const int nsq = s < N_SQUARINGS_CAP ? s : N_SQUARINGS_CAP;
  // squarings in pairs, A -> A2 -> A, so no copy between them
  int q = 0;
  for (; q + 2 <= nsq; q += 2) {
    quad_matmul6(A, A, A2);
    quad_matmul6(A2, A2, A);
  }
  if (q < nsq) {
    quad_matmul6(A, A, A2);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j) A[i][j] = A2[i][j];
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    double rs = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) rs += A[i][j];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double v = A[i][j] / rs;
      A[i][j] = isnan(v) ? 1.0 / 6.0 : v;
    }
  }
}


// ---------------------------------------------------------------------------
// The same expm by the two lanes of a pair (lanes 2m, 2m+1 of a quad): lane
// h = lane & 1 owns rows 3h..3h+2.  Row k of a product's right operand comes
// from lane k / 3 of the pair by one DPP quad permutation ([0,0,2,2] or
// [1,1,3,3]; the owner reads its own value through it), so every row again
// sees expm_ref's operations in expm_ref's order and the result is
// expm_ref's bit for bit up to the sign of a zero entry (as the quad's).
// Against the quad form a key costs 2 lanes instead of 4
// (the quad's fourth lane repeats the third's rows) and a product 72 DPP
// moves and 108 FMAs per lane instead of 72 and 72: half as many
// sub-rounds, each about a quarter longer, for the price of 36 VGPRs per
// matrix.  Both lanes of the pair must be active.
// ---------------------------------------------------------------------------
template <int H>
__device__ __forceinline__ double pair_bcast_d(double v) {
  constexpr int C = H == 0 ? 0xA0 : 0xF5;  // quad_perm [0,0,2,2] / [1,1,3,3]
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), C, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), C, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

template <int K, bool SPB>
__device__ __forceinline__ void pair_fma_row(const double A[3][6], const double B[3][6], double acc[3][6]) {
  double b[6];
#pragma unroll
  for (int j = 0; j < 6; ++j)
    if (!SPB || xb_gen_nz(K, j)) b[j] = pair_bcast_d<K / 3>(B[K % 3][j]);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      if (SPB && !xb_gen_nz(K, j)) {
        if (K == 0) acc[i][j] = 0.0;
      } else {
        acc[i][j] = K == 0 ? A[i][0] * b[j] : fma(A[i][K], b[j], acc[i][j]);
      }
    }
  __builtin_amdgcn_sched_barrier(0);
}
// C = A B for the lane's three rows; C must not alias A or B.  SPB: B has
// the XB generator's zero pattern.
template <bool SPB = false>
__device__ __forceinline__ void pair_matmul6(const double A[3][6], const double B[3][6], double C[3][6]) {
  pair_fma_row<0, SPB>(A, B, C);
  pair_fma_row<1, SPB>(A, B, C);
  pair_fma_row<2, SPB>(A, B, C);
  pair_fma_row<3, SPB>(A, B, C);
  pair_fma_row<4, SPB>(A, B, C);
  pair_fma_row<5, SPB>(A, B, C);
}

// A <- expm(A) for a 6x6 matrix held as row triples by a pair; h = lane & 1.
__device__ inline void expm_pair6(double A[3][6], int h) {
  double nrm = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double rs = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) rs += fabs(A[i][j]);
    nrm = fmax(nrm, rs);
  }
  nrm = fmax(pair_bcast_d<0>(nrm), pair_bcast_d<1>(nrm));
  int s = (int)ceilf(log2f((float)nrm / 0.5f + 1e-10f));
  s = s < 0 ? 0 : s;
  const double sc = ldexp(1.0, -s);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) A[i][j] *= sc;
  const double b0 = 1.0, b1 = 1.0 / 2.0, b2 = 1.0 / 9.0, b3 = 1.0 / 72.0, b4 = 1.0 / 1008.0,
               b5 = 1.0 / 30240.0, b6 = 1.0 / 1209600.0;
  double A2[3][6], A4[3][6], A6[3][6];
  pair_matmul6<true>(A, A, A2);
  pair_matmul6(A2, A2, A4);
  pair_matmul6(A4, A2, A6);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double id = (3 * h + i == j) ? 1.0 : 0.0;
      const double v = b0 * id + b2 * A2[i][j] + b4 * A4[i][j] + b6 * A6[i][j];
      const double w = b1 * id + b3 * A2[i][j] + b5 * A4[i][j];
      A6[i][j] = v;
      A4[i][j] = w;
    }
  double (&V)[3][6] = A6;
  double (&W)[3][6] = A4;
  pair_matmul6<true>(W, A, A2);  // U = W A
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      W[i][j] = V[i][j] - A2[i][j];
      A[i][j] = V[i][j] + A2[i][j];
    }
  // Gauss-Jordan as in expm_quad6 (pivot rows unscaled until the end)
  double ipiv[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double pw[6], pa[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      pw[j] = k < 3 ? pair_bcast_d<0>(W[k % 3][j]) : pair_bcast_d<1>(W[k % 3][j]);
      pa[j] = k < 3 ? pair_bcast_d<0>(A[k % 3][j]) : pair_bcast_d<1>(A[k % 3][j]);
    }
    const double inv = 1.0 / pw[k];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const bool own = 3 * h + i == k;
      ipiv[i] = own ? inv : ipiv[i];
      const double f = own ? 0.0 : W[i][k] * inv;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        W[i][j] = fma(-f, pw[j], W[i][j]);
        A[i][j] = fma(-f, pa[j], A[i][j]);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) A[i][j] *= ipiv[i];
  const int nsq = s < N_SQUARINGS_CAP ? s : N_SQUARINGS_CAP;
  int q = 0;
  for (; q + 2 <= nsq; q += 2) {
    pair_matmul6(A, A, A2);
    pair_matmul6(A2, A2, A);
  }
  if (q < nsq) {
    pair_matmul6(A, A, A2);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j) A[i][j] = A2[i][j];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double rs = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) rs += A[i][j];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double v = A[i][j] / rs;
      A[i][j] = isnan(v) ? 1.0 / 6.0 : v;
    }
  }
}

}  // namespace mf
