/*
 * cpu_port.c -- TEST INFRASTRUCTURE ONLY: a C restatement of the reference's
 * per-timestep path, used as the timed CPU baseline and as a second,
 * independently written check of the NumPy oracle.  Only tests/ and
 * bench.py's cpu_baseline leg load it.
 *
 * Arithmetic is FP32 with the reference solver's caps (solver.py:369-427,
 * 714-718): Newton-CG with tol = max(solver_tol, thick_k 1e-4, 0.25), at
 * most n_newton Newton and n_cg unrolled PCG iterations, the backbone
 * tridiagonal preconditioner applied as two prefix sums (its exact inverse,
 * state.py:269-316).  Instances run in parallel with OpenMP.
 *
 * Per step (simulation.py:679-721, timestep.py:65-199):
 *   thin += dz; nearest binding site per head (geometry.py:60-262);
 *   TM transitions with the bound-open lock (transitions.py:516-625);
 *   XB transitions and binding bookkeeping (transitions.py:1081-1245), the
 *   duplicate-bind rule being "clears first, lowest head index owns the
 *   site"; equilibrium solve; M-line force in double (forces.py:588-627).
 * Uniforms: Philox4x32-10, key (seed lo, seed hi), counter (i / 4, step,
 * purpose, 0), u = (word >> 9) 2^-23 -- the device's stream.
 * Transition tables come from the caller as float32 cumulative rows.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/multifil_amd.h"

typedef struct {
  float thick_k, thin_k, titin_a, titin_b, titin_rest;
  float g_k_w, g_rest_w, c_k_w, c_rest_w, g_k_s, g_rest_s, c_k_s, c_rest_s;
  float solver_tol, z_line0, ls0;
} mfc_params;

/* ---------------------------------------------------------------- Philox */
static void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)c[0] * 0xD2511F53u, p1 = (uint64_t)c[2] * 0xCD9E8D57u;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1, n3 = (uint32_t)p0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}
static void uniforms(uint64_t seed, int step, int purpose, int n, float* u) {
  for (int b = 0; 4 * b < n; ++b) {
    uint32_t c[4] = {(uint32_t)b, (uint32_t)step, (uint32_t)purpose, 0u};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    for (int k = 0; k < 4 && 4 * b + k < n; ++k) u[4 * b + k] = (float)(c[k] >> 9) * (1.0f / 8388608.0f);
  }
}

/* --------------------------------------------------------------- lattice */
typedef struct {
  int Tk, nc, Tn, ns, nx, nf, W, nb, nk, N, X, S;
  const float *co, *cr, *bo, *br, *edges;
  const int32_t *xthin, *xslot, *xvalid, *fsites, *fcount, *tprev, *tnext;
} Lat;

static Lat lattice(const mf_lattice* G, const float* lf, const int32_t* li) {
  Lat L;
  L.Tk = G->n_thick; L.nc = G->n_crowns; L.Tn = G->n_thin; L.ns = G->n_sites;
  L.nx = G->n_xb_per_crown; L.nf = G->n_faces; L.W = G->width; L.nb = G->n_bins;
  L.nk = L.Tk * L.nc; L.N = L.nk + L.Tn * L.ns; L.X = L.nk * L.nx; L.S = L.Tn * L.ns;
  L.co = lf + G->crown_offsets; L.cr = lf + G->crown_rests; L.bo = lf + G->binding_offsets;
  L.br = lf + G->binding_rests; L.edges = lf + G->bin_edges;
  L.xthin = li + G->xb_thin; L.xslot = li + G->xb_slot; L.xvalid = li + G->xb_valid;
  L.fsites = li + G->face_sites; L.fcount = li + G->face_count;
  L.tprev = li + G->tm_prev; L.tnext = li + G->tm_next;
  return L;
}

/* ------------------------------------------------------------- mechanics */
typedef struct {
  float *x, *f, *p, *r, *z, *Ap, *dx, *kx, *ktip;
  int* site;            /* [X] bound site DOF or -1 */
  unsigned char* st;    /* [X] XB state */
} Mech;

static void springs(const mfc_params* P, int st, float* gk, float* gr, float* ck, float* cr) {
  const int strong = st == 2 || st == 3;
  *gk = strong ? P->g_k_s : P->g_k_w; *gr = strong ? P->g_rest_s : P->g_rest_w;
  *ck = strong ? P->c_k_s : P->c_k_w; *cr = strong ? P->c_rest_s : P->c_rest_w;
}

/* residual F(x); refreshes kx / ktip; returns max|F| */
static float residual(const Lat* L, const mfc_params* P, Mech* m, float zl, float d) {
  const float kk = P->thick_k, kn = P->thin_k;
  for (int t = 0; t < L->Tk; ++t)
    for (int i = 0; i < L->nc; ++i) {
      const int n = t * L->nc + i;
      const float xm = i > 0 ? m->x[n - 1] : 0.0f;
      const float lo = (m->x[n] - xm - L->cr[n]) * kk;
      float hi;
      if (i < L->nc - 1) {
        hi = (m->x[n + 1] - m->x[n] - L->cr[n + 1]) * kk;
      } else {  /* titin on the tip crown (forces.py:136-157) */
        const float ad = zl - m->x[n], Lq = sqrtf(ad * ad + d * d);
        const float arg = P->titin_b * (Lq - P->titin_rest);
        const float ac = arg < -100.0f ? -100.0f : (arg > 100.0f ? 100.0f : arg);
        float F = P->titin_a * expf(ac);
        F = F > 0.0f ? F : 0.0f;
        const float Fp = (arg > -100.0f && arg < 100.0f) ? P->titin_b * F : 0.0f;
        hi = 6.0f * F * (Lq > 0.0f ? ad / Lq : 0.0f);
        m->ktip[t] = -6.0f * (Fp * ad * ad / (Lq * Lq) + F * d * d / (Lq * Lq * Lq));
      }
      m->f[n] = hi - lo;
    }
  for (int g = 0; g < L->Tn; ++g)
    for (int i = 0; i < L->ns; ++i) {
      const int n = L->nk + g * L->ns + i, k = g * L->ns + i;
      const float xp = i < L->ns - 1 ? m->x[n + 1] : zl;
      const float hi = (xp - m->x[n] - L->br[k]) * kn;
      const float lo = i > 0 ? (m->x[n] - m->x[n - 1] - L->br[k - 1]) * kn : 0.0f;
      m->f[n] = hi - lo;
    }
  for (int j = 0; j < L->X; ++j) {  /* XB forces (forces.py:277-418) */
    m->kx[j] = 0.0f;
    const int s = m->site[j], st = m->st[j];
    if (s < 0 || st < 1 || st > 3) continue;
    const int c = j / L->nx;
    float gk, gr, ck, cr;
    springs(P, st, &gk, &gr, &ck, &cr);
    const float xd = m->x[s] - m->x[c];
    const float r = sqrtf(xd * xd + d * d), rs = r > 1e-10f ? r : 1e-10f;
    const float th = atan2f(d, xd);
    const float f = gk * (r - gr) * (xd / rs) - (1.0f / rs) * ck * (th - cr) * (d / rs);
    m->kx[j] = gk * (1.0f - gr * d * d / (rs * rs * rs)) + ck * d * (d + 2.0f * xd * (th - cr)) / (rs * rs * rs * rs);
    m->f[c] += f;
    m->f[s] -= f;
  }
  float mx = 0.0f;
  for (int n = 0; n < L->N; ++n) mx = fmaxf(mx, fabsf(m->f[n]));
  return mx;
}

/* Ap = -J p */
static void neg_jvp(const Lat* L, const mfc_params* P, const Mech* m, const float* p, float* Ap) {
  const float kk = P->thick_k, kn = P->thin_k;
  for (int t = 0; t < L->Tk; ++t)
    for (int i = 0; i < L->nc; ++i) {
      const int n = t * L->nc + i;
      const float pm = i > 0 ? p[n - 1] : 0.0f;
      const float lo = kk * (p[n] - pm);
      const float hi = i < L->nc - 1 ? kk * (p[n + 1] - p[n]) : m->ktip[t] * p[n];
      Ap[n] = -(hi - lo);
    }
  for (int g = 0; g < L->Tn; ++g)
    for (int i = 0; i < L->ns; ++i) {
      const int n = L->nk + g * L->ns + i;
      const float pp = i < L->ns - 1 ? p[n + 1] : 0.0f;
      const float hi = kn * (pp - p[n]);
      const float lo = i > 0 ? kn * (p[n] - p[n - 1]) : 0.0f;
      Ap[n] = -(hi - lo);
    }
  for (int j = 0; j < L->X; ++j) {
    if (m->kx[j] == 0.0f) continue;
    const int s = m->site[j], c = j / L->nx;
    const float q = m->kx[j] * (p[s] - p[c]);
    Ap[c] -= q;
    Ap[s] += q;
  }
}

/* z = M^-1 r: thick chains prefix(suffix(r))/k, thin chains suffix(prefix(r))/k */
static void precond(const Lat* L, const mfc_params* P, const float* r, float* z) {
  for (int t = 0; t < L->Tk; ++t) {
    const float* rr = r + t * L->nc;
    float* zz = z + t * L->nc;
    float acc = 0.0f;
    for (int i = L->nc - 1; i >= 0; --i) { acc += rr[i]; zz[i] = acc; }
    acc = 0.0f;
    for (int i = 0; i < L->nc; ++i) { acc += zz[i]; zz[i] = acc / P->thick_k; }
  }
  for (int g = 0; g < L->Tn; ++g) {
    const float* rr = r + L->nk + g * L->ns;
    float* zz = z + L->nk + g * L->ns;
    float acc = 0.0f;
    for (int i = 0; i < L->ns; ++i) { acc += rr[i]; zz[i] = acc; }
    acc = 0.0f;
    for (int i = L->ns - 1; i >= 0; --i) { acc += zz[i]; zz[i] = acc / P->thin_k; }
  }
}

/* Newton-CG (solver.py:329-427, 663-771); returns max|F|, *iters */
static float solve(const Lat* L, const mfc_params* P, Mech* m, float zl, float d, int n_newton, int n_cg,
                   int* iters) {
  const float tol = fmaxf(P->solver_tol, fmaxf(P->thick_k * 1e-4f, 0.25f));
  const int N = L->N;
  float mx = residual(L, P, m, zl, d);
  int it = 0;
  while (mx > tol && it < n_newton) {
    for (int n = 0; n < N; ++n) { m->r[n] = m->f[n]; m->dx[n] = 0.0f; m->p[n] = 0.0f; }
    float rz = 0.0f;
    for (int k = 0; k < n_cg; ++k) {
      precond(L, P, m->r, m->z);
      float rzn = 0.0f;
      for (int n = 0; n < N; ++n) rzn += m->r[n] * m->z[n];
      const float beta = k == 0 ? 0.0f : rzn / (rz + 1e-30f);
      rz = rzn;
      for (int n = 0; n < N; ++n) m->p[n] = m->z[n] + beta * m->p[n];
      neg_jvp(L, P, m, m->p, m->Ap);
      float pAp = 0.0f;
      for (int n = 0; n < N; ++n) pAp += m->p[n] * m->Ap[n];
      const float alpha = rz / (pAp + 1e-30f);
      for (int n = 0; n < N; ++n) { m->dx[n] += alpha * m->p[n]; m->r[n] -= alpha * m->Ap[n]; }
    }
    for (int n = 0; n < N; ++n) m->x[n] += isfinite(m->dx[n]) ? m->dx[n] : 0.0f;
    mx = residual(L, P, m, zl, d);
    ++it;
  }
  *iters = it;
  return mx;
}

/* ------------------------------------------------------------- kinetics */
static void nearest(const Lat* L, const float* x, int j, int* near, float* xd) {
  const int c = j / L->nx, slot = L->xslot[j], thin = L->xthin[j];
  const float base = x[c], head = base + 13.0f;
  const int cnt = L->fcount[slot], tb = L->nk + thin * L->ns;
  const int32_t* cs = L->fsites + slot * L->W;
  float best = INFINITY;
  int bj = 0;
  for (int k = 0; k < L->W; ++k) {
    const float sp = x[tb + cs[k]];
    if (k < cnt && sp > 0.0f && fabsf(sp - head) < best) { best = fabsf(sp - head); bj = k; }
  }
  *near = cs[bj];
  *xd = x[tb + cs[bj]] - base;
}

static int bin_of(const Lat* L, float x) {  /* clip(digitize(x, edges) - 1, 0, nb - 1) */
  if (isnan(x)) return L->nb - 1;
  int lo = 0, hi = L->nb + 1;  /* count of edges <= x */
  while (lo < hi) { const int mid = (lo + hi) / 2; if (L->edges[mid] <= x) lo = mid + 1; else hi = mid; }
  const int b = lo - 1;
  return b < 0 ? 0 : (b > L->nb - 1 ? L->nb - 1 : b);
}

static int sample(const float* cum, int n, float u) {
  for (int k = 0; k < n; ++k)
    if (u < cum[k]) return k;
  return 0;
}

/* one instance, T steps */
static void run_one(const Lat* L, const mfc_params* P, const float* tm_cum, const float* xb_cum,
                    const float* drv, int T, uint64_t seed, int n_newton, int n_cg, double* force,
                    float* resid, int* iters, int* nbound) {
  const int N = L->N, X = L->X, S = L->S;
  Mech m;
  float* buf = (float*)calloc((size_t)8 * N + 2 * X + L->Tk + S + X, sizeof(float));
  m.x = buf; m.f = buf + N; m.p = buf + 2 * N; m.r = buf + 3 * N; m.z = buf + 4 * N; m.Ap = buf + 5 * N;
  m.dx = buf + 6 * N; m.kx = buf + 8 * N; m.ktip = m.kx + X;
  float* u = m.ktip + L->Tk;  /* [S + X] draws */
  m.site = (int*)calloc(X, sizeof(int));
  m.st = (unsigned char*)calloc(X, 1);
  int* tbound = (int*)malloc(sizeof(int) * S);
  unsigned char* tm = (unsigned char*)calloc(S, 1);
  unsigned char* tmn = (unsigned char*)calloc(S, 1);
  int* near = (int*)malloc(sizeof(int) * X);
  int* occ = (int*)malloc(sizeof(int) * X);
  float* xd = (float*)malloc(sizeof(float) * X);
  /* realize_state + equilibrate (simulation.py:633-642), solved at the
   * parameter row's z_line / spacing */
  for (int n = 0; n < N; ++n) m.x[n] = n < L->nk ? L->co[n] : drv[0] - L->bo[n - L->nk];
  for (int j = 0; j < X; ++j) m.site[j] = -1;
  for (int s = 0; s < S; ++s) tbound[s] = -1;
  int it;
  solve(L, P, &m, P->z_line0, P->ls0, n_newton, n_cg, &it);
  for (int t = 0; t < T; ++t) {
    const float zl = drv[3 * t], d = drv[3 * t + 2];
    const float dz = t > 0 ? zl - drv[3 * (t - 1)] : 0.0f;
    for (int n = L->nk; n < N; ++n) m.x[n] += dz;
    uniforms(seed, t, 0, S, u);
    uniforms(seed, t, 1, X, u + S);
    for (int j = 0; j < X; ++j) nearest(L, m.x, j, &near[j], &xd[j]);
    for (int s = 0; s < S; ++s) {  /* TM transitions */
      const int row = (s / L->ns) * L->ns, i = s - row;
      const int pv = L->tprev[s], nv = L->tnext[s];
      const int sp = tm[row + pv], sn = tm[row + nv], pr = pv != i, nr = nv != i;
      const int n2 = (pr && sp == 2) + (nr && sn == 2), n3 = (pr && sp == 3) + (nr && sn == 3);
      const int ncl = (pr && sp <= 1) + (nr && sn <= 1);
      const int cfg = 9 * n2 + 3 * n3 + ncl;
      tmn[s] = (tm[s] == 3 && tbound[s] >= 0) ? 3 : (unsigned char)sample(tm_cum + cfg * 16 + tm[s] * 4, 4, u[s]);
    }
    for (int j = 0; j < X; ++j) occ[j] = tbound[L->xthin[j] * L->ns + near[j]] >= 0;
    for (int j = 0; j < X; ++j) {  /* XB transitions; clears before binds */
      const int old = m.st[j], gs = L->xthin[j] * L->ns + near[j];
      const int perm = tmn[gs] == 3 && L->xvalid[j];
      const int key = (perm ? L->nb : 0) + bin_of(L, xd[j]);
      int nw = sample(xb_cum + key * 36 + old * 6, 6, u[S + j]);
      const int ob = old >= 1 && old <= 3, nbd = nw >= 1 && nw <= 3;
      if (ob && !nbd) {
        if (m.site[j] >= 0) tbound[m.site[j] - L->nk] = -1;
        m.site[j] = -1;
      }
      const int was_occ = occ[j];
      occ[j] = 0;  /* from here: "bound this step" */
      if (!ob && nbd) {
        if (perm && !was_occ) { m.site[j] = L->nk + gs; occ[j] = 1; }
        else nw = 0;
      }
      m.st[j] = (unsigned char)nw;
    }
    for (int j = 0; j < X; ++j) {  /* of the heads bound this step, the lowest index owns the site */
      const int s = m.site[j];
      if (occ[j] && (tbound[s - L->nk] < 0 || tbound[s - L->nk] > j)) tbound[s - L->nk] = j;
    }
    memcpy(tm, tmn, S);
    resid[t] = solve(L, P, &m, zl, d, n_newton, n_cg, &iters[t]);
    double fs = 0.0;
    for (int k = 0; k < L->Tk; ++k) fs += ((double)m.x[k * L->nc] - (double)L->co[k * L->nc]) * (double)P->thick_k;
    force[t] = fs;
    int nbd = 0;
    for (int j = 0; j < X; ++j) nbd += m.st[j] >= 1 && m.st[j] <= 3;
    nbound[t] = nbd;
  }
  free(buf); free(m.site); free(m.st); free(tbound); free(tm); free(tmn); free(near); free(occ); free(xd);
}

/* B instances of one lattice, T steps each; tables per instance. */
int mfc_run(const mf_lattice* G, const float* lf, const int32_t* li, const mfc_params* P,
            const float* tm_cum, const float* xb_cum, int B, int T, const float* drivers,
            const uint64_t* seeds, int n_newton, int n_cg, int n_threads, double* force, float* resid,
            int* iters, int* nbound) {
  const Lat L = lattice(G, lf, li);
  const int ntab_tm = 27 * 16, ntab_xb = 2 * L.nb * 36;
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads)
  for (int b = 0; b < B; ++b)
    run_one(&L, P + b, tm_cum + (size_t)b * ntab_tm, xb_cum + (size_t)b * ntab_xb, drivers + (size_t)b * T * 3, T,
            seeds[b], n_newton, n_cg, force + (size_t)b * T, resid + (size_t)b * T, iters + (size_t)b * T,
            nbound + (size_t)b * T);
  return 0;
}
