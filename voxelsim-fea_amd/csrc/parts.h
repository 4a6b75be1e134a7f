// parts.h -- many parts of uneven size in one launch per operation
// (vx_parts_* / vx_solve_parts).  Included at the end of vxfea.hip: it
// reuses the single-grid solver's helpers, step coefficients and kernel
// bodies.
//
// The reference solves every part with its own VoxelFEA / MGSolver
// (core/fea.py:244-305, core/multigrid.py:439-583).  Here each part keeps
// its own grid, hierarchy (the rule of core/multigrid.py:141-148), fixed
// DOFs, density and PCG recurrence (its own alpha, beta, ratio, stall guard
// and freeze, as vx_solve_batch does for load cases), while every
// operator, smoother, transfer and reduction launch covers all the parts of
// a level: the launches per iteration no longer grow with the part count,
// and a converged part's work-groups return at once.
//
// Layout.  Parts are ordered by hierarchy depth (deepest first; among the
// parts whose coarsest level is the same, those small enough for the one-
// work-group coarse kernel first).  At level l the parts present are then a
// prefix [0, np) of that order, of which [0, m) go on to level l+1 and
// [m, np) end there: [m, mc) take the one-work-group coarsest smoothing,
// [mc, np) tile-kernel sweeps.  A level's vectors, node arrays and element
// arrays pack its parts one after another, each part's block starting on a
// 64-element boundary.
#pragma once

namespace {

constexpr long long PART_ALIGN = 64;

long long align_up(long long v) { return (v + PART_ALIGN - 1) / PART_ALIGN * PART_ALIGN; }

struct PLevel {
  int np = 0, m = 0, mc = 0;
  long long n3 = 0, nn = 0, ne = 0;   // packed values, nodes, elements
  std::vector<PartGrid> hg;
  PartGrid* dg = nullptr;
  std::vector<PartTiles> ht;          // np + 1 entries (the last holds the total)
  PartTiles* dt = nullptr;
  std::vector<PartTr> hr, hp;         // restriction to / prolongation from level l+1, m + 1
  PartTr* dr = nullptr;
  PartTr* dp = nullptr;
  int coarse_threads = 64;            // work-group of the one-work-group coarse kernel
  uint8_t *mask = nullptr, *mask_fixed = nullptr;
  double *E64 = nullptr, *mi64 = nullptr;
  float *E32 = nullptr, *mi32 = nullptr;
  double *b64 = nullptr, *x64 = nullptr, *t64 = nullptr, *w64 = nullptr;
  float *b32 = nullptr, *x32 = nullptr, *t32 = nullptr, *w32 = nullptr;
};

template <typename T> struct PSel;
template <> struct PSel<double> {
  static double* b(PLevel& l) { return l.b64; }
  static double* x(PLevel& l) { return l.x64; }
  static double* t(PLevel& l) { return l.t64; }
  static double* w(PLevel& l) { return l.w64; }
  static double* E(PLevel& l) { return l.E64; }
  static double* mi(PLevel& l) { return l.mi64; }
};
template <> struct PSel<float> {
  static float* b(PLevel& l) { return l.b32; }
  static float* x(PLevel& l) { return l.x32; }
  static float* t(PLevel& l) { return l.t32; }
  static float* w(PLevel& l) { return l.w32; }
  static float* E(PLevel& l) { return l.E32; }
  static float* mi(PLevel& l) { return l.mi32; }
};

}  // namespace

struct vx_parts {
  vx_solver s;                        // options, element constants, stream, scalars
                                      // ([np]), partials, allocations, last error
  int np = 0;
  std::vector<std::array<int, 3>> dims;   // level-0 element dims, caller order
  std::vector<int> nlev;                  // levels per part, caller order
  std::vector<int> order;                 // internal position -> caller index
  std::vector<long long> uoff, ueoff;     // caller packing: values, elements (caller order)
  long long u_total = 0, e_total = 0;
  std::vector<PLevel> lv;
  double *u64 = nullptr, *r64 = nullptr, *z64 = nullptr, *p64 = nullptr, *Ap64 = nullptr,
         *fm64 = nullptr;
  float *r32 = nullptr, *z32 = nullptr, *p32 = nullptr, *Ap32 = nullptr, *d32 = nullptr;
  bool fixed_set = false, density_set = false;
};

namespace {

// Where a smoothing pass runs for the parts [q0, q1) of level l; fz (may be
// null): skip frozen parts.
struct PartsTarget {
  vx_parts* P;
  int l, q0, q1;
  const Scalars* fz;
};

vx_solver* tsol(const PartsTarget& t) { return &t.P->s; }

long long tfirst(const PartsTarget& t) { return t.P->lv[t.l].hg[t.q0].voff; }

long long tvals(const PartsTarget& t) {
  const PLevel& V = t.P->lv[t.l];
  const PartGrid& g = V.hg[t.q1 - 1];
  return g.voff + 3LL * (g.nx + 1) * (g.ny + 1) * (g.nz + 1) - tfirst(t);
}

template <typename T, int MODE, bool DOT>
int ttile(const PartsTarget& t, const T* u, const T* b, T* out, int* n_part,
          const SmoothCoef* sc) {
  vx_solver* s = &t.P->s;
  PLevel& V = t.P->lv[t.l];
  TileArgs<T> a{};
  a.nb = 1;
  a.u = u; a.E = PSel<T>::E(V); a.mask = V.mask; a.b = b; a.out = out;
  a.omega = (T)(sc ? sc->omega : s->opt.omega);
  a.up = sc ? static_cast<const T*>(sc->up) : nullptr;
  a.cu = (T)(sc ? sc->cu : 0.0);
  a.cr = (T)(sc ? sc->cr : 0.0);
  a.kd = (T)s->kd;
  a.K = kconst<T>(s);
  a.partial = DOT ? s->partial : nullptr;
  const long long total = V.ht[t.q1].begin - V.ht[t.q0].begin;
  if (n_part) *n_part = (int)total;
  const unsigned launch = (unsigned)((total + 7) / 8 * 8);   // XCD-ordered 1D grid
  hipLaunchKernelGGL((k_tile_parts<T, MODE, DOT>), dim3(launch), dim3(NT), 0, s->stream, a, V.dg,
                     V.dt, t.q0, t.q1, t.fz);
  LAUNCHCHK();
  return VX_OK;
}

template <typename T>
int tzero(const PartsTarget& t, const T* b, T* out, double omega, bool dot, int*) {
  vx_solver* s = &t.P->s;
  if (dot) return fail(s, VX_ERR_ARG, "parts: a one-step smoothing pass with dot products");
  PLevel& V = t.P->lv[t.l];
  long long nmax = 1;
  for (int q = t.q0; q < t.q1; ++q)
    nmax = std::max<long long>(nmax, (long long)(V.hg[q].nx + 1) * (V.hg[q].ny + 1) *
                                         (V.hg[q].nz + 1));
  const bool aligned = ((reinterpret_cast<uintptr_t>(b) | reinterpret_cast<uintptr_t>(out) |
                         reinterpret_cast<uintptr_t>(PSel<T>::mi(V)) |
                         reinterpret_cast<uintptr_t>(V.mask)) & 15) == 0;
  hipLaunchKernelGGL(k_zero_step_parts<T>, dim3(ew_grid(nmax), t.q1 - t.q0), dim3(EW_THREADS), 0,
                     s->stream, V.dg, t.q0, V.mask, PSel<T>::mi(V), b, out, (T)omega,
                     aligned ? 1 : 0, t.fz);
  LAUNCHCHK();
  return VX_OK;
}

// per part: sum its partials and apply op to its scalars; tiles: the level-0
// tile table gives each part's partials, else EW_BLOCKS per part
int pfinalize(vx_parts* P, bool tiles, int op) {
  vx_solver* s = &P->s;
  hipLaunchKernelGGL(k_finalize_parts, dim3(P->np), dim3(FIN_THREADS), 0, s->stream, s->partial,
                     tiles ? P->lv[0].dt : nullptr, EW_BLOCKS, op, s->dscal);
  LAUNCHCHK();
  return VX_OK;
}

// V-cycle of every part (core/multigrid.py:426-436) from level 0: the parts
// that end at a level take their coarsest smoothing there; the others
// pre-smooth, restrict, and after the coarser levels prolong and
// post-smooth.  dot: partials of z0.b0 (level-0 tile layout).
template <typename T>
int pvcycle(vx_parts* P, const T* b0, T* z0, T* t0, T* w0, bool dot, bool first_done) {
  vx_solver* s = &P->s;
  const int L = (int)P->lv.size();
  const Scalars* fz = s->dscal;
  bool restricted_x1 = false;
  for (int l = 0; l < L; ++l) {
    PLevel& V = P->lv[l];
    const T* b = l == 0 ? b0 : PSel<T>::b(V);
    T* x = l == 0 ? z0 : PSel<T>::x(V);
    T* t = l == 0 ? t0 : PSel<T>::t(V);
    T* w = l == 0 ? w0 : PSel<T>::w(V);
    if (V.np > V.m) {
      // parts whose coarsest level this is: coarse_sweeps steps from zero
      const int k = s->opt.coarse_sweeps;
      if (V.mc > V.m) {
        const CoarseCoef cc = smooth_coef(s, k);
        hipLaunchKernelGGL(k_coarse_smooth_parts<T>, dim3(V.mc - V.m), dim3(V.coarse_threads), 0,
                           s->stream, V.dg, V.m, PSel<T>::E(V), V.mask, PSel<T>::mi(V), b, x,
                           kconst<T>(s), cc, fz);
        LAUNCHCHK();
      }
      if (V.np > V.mc)
        CHK(tsmooth_step<T>(PartsTarget{P, l, V.mc, V.np, fz}, nullptr, b, k, x, t, w,
                            dot && l == 0, nullptr));
    }
    if (V.m == 0) break;
    const PartsTarget go{P, l, 0, V.m, fz};
    const int ns = level_ns(s, l);
    CHK(tsmooth_step<T>(go, nullptr, b, ns, x, t, w, false, nullptr,
                        l == 0 ? first_done : restricted_x1));
    CHK((ttile<T, M_RESID, false>(go, x, b, t, nullptr, nullptr)));
    // restriction into level l+1, whose parts are [0, V.m); those that smooth
    // there from zero (not their coarsest) also get their first step
    PLevel& C = P->lv[l + 1];
    const int nsc = level_ns(s, l + 1);
    restricted_x1 = C.m > 0 && first_in_tmp(s, nsc);
    const unsigned blocks = (unsigned)((V.hr[V.m].begin + 7) / 8 * 8);
    hipLaunchKernelGGL(k_restrict_parts<T>, dim3(blocks), dim3(TR_THREADS), 0, s->stream, V.dr, 0,
                       V.m, t, C.mask, PSel<T>::b(C), PSel<T>::mi(C), (T)first_step_coef(s, nsc),
                       restricted_x1 ? PSel<T>::t(C) : nullptr, C.m, fz);
    LAUNCHCHK();
  }
  for (int l = L - 2; l >= 0; --l) {
    PLevel& V = P->lv[l];
    if (V.m == 0) continue;
    PLevel& C = P->lv[l + 1];
    const T* b = l == 0 ? b0 : PSel<T>::b(V);
    T* x = l == 0 ? z0 : PSel<T>::x(V);
    T* t = l == 0 ? t0 : PSel<T>::t(V);
    T* w = l == 0 ? w0 : PSel<T>::w(V);
    const unsigned blocks = (unsigned)((V.hp[V.m].begin + 7) / 8 * 8);
    hipLaunchKernelGGL(k_prolong_parts<T>, dim3(blocks), dim3(TR_THREADS), 0, s->stream, V.dp, 0,
                       V.m, (const T*)PSel<T>::x(C), V.mask, x, fz);
    LAUNCHCHK();
    CHK(tsmooth_step<T>(PartsTarget{P, l, 0, V.m, fz}, x, b, level_ns(s, l), x, t, w,
                        dot && l == 0, nullptr));
  }
  return VX_OK;
}

// PCG of every part at once (core/multigrid.py:474-546), each part with its
// own scalars: the recurrence of pcg() above, on the parts launches.  Parts
// flagged in `skip` start frozen; a part freezes at its convergence or
// stall.  The work-groups of frozen parts return at once, except the
// pending deferred update u += alpha p, which the next direction update
// still applies.
template <typename T>
int ppcg(vx_parts* P, T* u, T* r, T* z, T* p, T* Ap, T* t0, T* w0, const std::vector<double>& tol_rel, int max_iter,
         bool stall_guard, const std::vector<double>& ratio_scale, double cb_tol,
         const std::vector<int>& iter_offset, vx_progress_cb cb, void* user, double* hist,
         int hist_cap, std::vector<PcgResult>& res, const std::vector<int>& skip) {
  vx_solver* s = &P->s;
  PLevel& V0 = P->lv[0];
  const int nb = P->np;
  hipEvent_t ev;
  HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  auto cleanup = [&]() { (void)hipEventDestroy(ev); };
  res.assign(nb, PcgResult());
  std::vector<int> frozen(skip.begin(), skip.end());
  // frozen flags of an earlier solve must not skip this one's first V-cycle
  int rc = push_frozen(s, frozen);
  if (rc == VX_OK) rc = pvcycle<T>(P, r, z, t0, w0, true, false);
  if (rc == VX_OK) rc = pfinalize(P, true, F_RZ0);   // also beta = 0, frozen = 0
  if (rc == VX_OK) rc = push_frozen(s, frozen);
  if (rc != VX_OK) { cleanup(); return rc; }
  const auto al16 = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
  const bool fuse_x1 = std::is_same<T, float>::value && first_in_tmp(s, level_ns(s, 0)) &&
                       V0.m == V0.np && al16(u) && al16(r) && al16(p) && al16(Ap) && al16(t0) &&
                       al16(V0.mask) && al16(V0.mi32);
  const bool lazy_u = fuse_x1;
  const double x1_coef = first_step_coef(s, level_ns(s, 0));
  if (hipMemsetAsync(p, 0, sizeof(T) * V0.n3, s->stream) != hipSuccess) {
    cleanup();
    return fail(s, VX_ERR_HIP, "memset p");
  }
  const PartsTarget all{P, 0, 0, nb, s->dscal};
  bool graphable = !s->opt.eager;
  for (int l = 0; l < (int)P->lv.size() && graphable; ++l) {
    const int k = level_ns(s, l);   // smooth_step copies for these (copy_vec)
    if (s->opt.smoother == VX_SMOOTH_CHEBYSHEV ? k == 1 : k % 2 == 1) graphable = false;
  }
  std::vector<double> best(nb, INFINITY), prev(nb, NAN), prev2(nb, NAN);
  for (int it = 0; it < max_iter; ++it) {
    if (lazy_u)
      hipLaunchKernelGGL(k_p_update_u_parts<T>, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0,
                         s->stream, V0.dg, s->dscal, z, p, u);
    else
      hipLaunchKernelGGL(k_p_update_b_parts<T>, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0,
                         s->stream, V0.dg, s->dscal, z, p);
    if (hipGetLastError() != hipSuccess) { rc = fail(s, VX_ERR_HIP, "p update launch"); break; }
    rc = ttile<T, M_MATVEC, true>(all, p, nullptr, Ap, nullptr, nullptr);
    if (rc == VX_OK) rc = pfinalize(P, true, F_ALPHA);
    if (rc != VX_OK) break;
    if (fuse_x1)
      hipLaunchKernelGGL(k_cg_update_x1_parts, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0, s->stream,
                         V0.dg, s->dscal, reinterpret_cast<float*>(r),
                         reinterpret_cast<const float*>(Ap), s->partial, V0.mask, V0.mi32,
                         (float)x1_coef, reinterpret_cast<float*>(t0));
    else
      hipLaunchKernelGGL(k_cg_update_parts<T>, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0,
                         s->stream, V0.dg, s->dscal, u, r, p, Ap, s->partial);
    if (hipGetLastError() != hipSuccess) { rc = fail(s, VX_ERR_HIP, "cg_update launch"); break; }
    rc = pfinalize(P, false, F_RATIO);
    if (rc == VX_OK) rc = read_scalars(s, ev);
    if (rc != VX_OK) break;
    bool predicted_done = true;
    for (int q = 0; q < nb; ++q) {
      if (frozen[q]) continue;
      bool d = false;
      if (!std::isnan(prev[q]) && !std::isnan(prev2[q]) && prev2[q] > 0)
        d = prev[q] * (prev[q] / prev2[q]) < 3.0 * tol_rel[q];
      predicted_done = predicted_done && d;
    }
    // the V-cycle and its beta: the same launches every step, replayed from
    // a captured graph when no smoothing pass of the cycle holds a copy
    auto cycle_and_beta = [&]() -> int {
      CHK(pvcycle<T>(P, r, z, t0, w0, true, fuse_x1));
      return pfinalize(P, true, F_BETA);
    };
    auto queue_next = [&]() -> int {
      if (!graphable) return cycle_and_beta();
      const void* const key[4] = {r, z, t0, Ap};
      return run_graphed(s, key, std::is_same<T, float>::value ? 3 : 2, 0, fuse_x1 ? 1 : 0, 0,
                         cycle_and_beta);
    };
    bool queued = false;
    if (!predicted_done) {
      rc = queue_next();
      if (rc != VX_OK) break;
      queued = true;
    }
    if (hipEventSynchronize(ev) != hipSuccess) { rc = fail(s, VX_ERR_HIP, "event sync"); break; }
    bool changed = false;
    double worst = 0.0;
    for (int q = 0; q < nb; ++q) {
      if (frozen[q]) continue;
      const Scalars sc = s->hscal[q];
      PcgResult& R = res[q];
      if (sc.pAp == 0.0) {   // nothing left along p (core/multigrid.py:500-511)
        R.converged = R.ratio < tol_rel[q];
        frozen[q] = 1;
        changed = true;
        continue;
      }
      const double ratio = sc.ratio;
      R.iters = it + 1;
      R.ratio = ratio;
      const int gi = iter_offset[q] + it;
      if (hist && gi < hist_cap) hist[(long long)q * hist_cap + gi] = ratio * ratio_scale[q];
      worst = std::max(worst, ratio * ratio_scale[q]);
      if (ratio < tol_rel[q]) {
        R.converged = true;
        frozen[q] = 1;
        changed = true;
        continue;
      }
      best[q] = std::min(best[q], ratio);
      if (stall_guard && it + 1 >= 10 && ratio > 5.0 * best[q]) {
        R.stalled = true;
        frozen[q] = 1;
        changed = true;
        char buf[220];
        snprintf(buf, sizeof buf,
                 "MG-PCG stalled/diverged at iter %d (resid ratio %.3e, best seen %.3e) on part %d",
                 it + 1, ratio, best[q], P->order[q]);
        s->err = buf;
        continue;
      }
      prev2[q] = prev[q];
      prev[q] = ratio;
    }
    if (cb) cb(iter_offset[0] + it + 1, worst, cb_tol, user);
    if (std::all_of(frozen.begin(), frozen.end(), [](int f) { return f; })) break;
    if (changed) {
      rc = push_frozen(s, frozen);
      if (rc != VX_OK) break;
    }
    if (!queued) {
      rc = queue_next();
      if (rc != VX_OK) break;
    }
  }
  if (lazy_u && rc == VX_OK) {
    hipLaunchKernelGGL(k_u_flush_parts<T>, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0, s->stream,
                       V0.dg, s->dscal, p, u);
    if (hipGetLastError() != hipSuccess) rc = fail(s, VX_ERR_HIP, "u flush launch");
  }
  cleanup();
  return rc;
}

// z-chunks of the tile kernels for the parts of one level: options.z_chunk,
// else, as for one grid, halve the chunks (not below 16 planes) until the
// level's launch has at least 2048 work-groups in total
void part_zchunks(const vx_options& o, const std::vector<std::array<int, 3>>& d,
                  std::vector<int>& zc) {
  zc.resize(d.size());
  for (size_t q = 0; q < d.size(); ++q) zc[q] = o.z_chunk > 0 ? o.z_chunk : d[q][2] + 1;
  if (o.z_chunk > 0) return;
  for (;;) {
    long long total = 0;
    for (size_t q = 0; q < d.size(); ++q) {
      const long long cols =
          (long long)((d[q][0] + 1 + OWN - 1) / OWN) * ((d[q][1] + 1 + OWN - 1) / OWN);
      total += cols * ((d[q][2] + 1 + zc[q] - 1) / zc[q]);
    }
    if (total >= 2048) return;
    bool any = false;
    for (auto& z : zc)
      if (z > 16) {
        z = (z + 1) / 2;
        any = true;
      }
    if (!any) return;
  }
}

// a level's tile table (host) for per-part z-chunks zc; returns the total
// work-group count
int part_tiles(PLevel& V, const std::vector<int>& zc) {
  V.ht.resize(V.np + 1);
  int begin = 0;
  for (int q = 0; q < V.np; ++q) {
    const PartGrid& g = V.hg[q];
    PartTiles& t = V.ht[q];
    t.zc = zc[q];
    t.gx = (g.nx + 1 + OWN - 1) / OWN;
    t.gy = (g.ny + 1 + OWN - 1) / OWN;
    t.gz = (g.nz + 1 + zc[q] - 1) / zc[q];
    t.begin = begin;
    begin += t.gx * t.gy * t.gz;
  }
  V.ht[V.np] = PartTiles{0, 0, 0, 0, begin};
  return begin;
}

template <typename T>
int upload(vx_solver* s, T** dst, const std::vector<T>& v) {
  CHK(dalloc(s, dst, v.size()));
  HIPCHK(hipMemcpyAsync(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s->stream));
  return VX_OK;
}

// set-up of a parts handle after its options and per-part hierarchies are known
int parts_build(vx_parts* P) {
  vx_solver* s = &P->s;
  const int np = P->np;
  const bool f32 = s->opt.precision == VX_MIXED;
  std::vector<std::vector<std::array<int, 3>>> hier(np);
  int L = 0;
  for (int i = 0; i < np; ++i) {
    hier[i] = level_dims_rule(P->dims[i][0], P->dims[i][1], P->dims[i][2], s->opt);
    P->nlev[i] = (int)hier[i].size();
    L = std::max(L, P->nlev[i]);
  }
  // a part's coarsest level fits the one-work-group kernel
  auto small = [&](int i) {
    const auto& d = hier[i].back();
    const long long ne = (long long)d[0] * d[1] * d[2];
    const long long nn = (long long)(d[0] + 1) * (d[1] + 1) * (d[2] + 1);
    return P->nlev[i] > 1 && s->opt.coarse_sweeps <= COARSE_MAX_STEPS &&
           (f32 ? (ne <= coarse_max_elems<float>() && nn <= coarse_max_nodes<float>())
                : (ne <= coarse_max_elems<double>() && nn <= coarse_max_nodes<double>()));
  };
  P->order.resize(np);
  for (int i = 0; i < np; ++i) P->order[i] = i;
  std::stable_sort(P->order.begin(), P->order.end(), [&](int a, int b) {
    if (P->nlev[a] != P->nlev[b]) return P->nlev[a] > P->nlev[b];
    return small(a) && !small(b);
  });
  // caller packing
  P->uoff.resize(np);
  P->ueoff.resize(np);
  P->u_total = P->e_total = 0;
  for (int i = 0; i < np; ++i) {
    const auto& d = P->dims[i];
    P->uoff[i] = P->u_total;
    P->ueoff[i] = P->e_total;
    P->u_total += 3LL * (d[0] + 1) * (d[1] + 1) * (d[2] + 1);
    P->e_total += (long long)d[0] * d[1] * d[2];
  }
  P->lv.assign(L, PLevel());
  long long max_tiles = 0;
  for (int l = 0; l < L; ++l) {
    PLevel& V = P->lv[l];
    std::vector<std::array<int, 3>> d;
    for (int q = 0; q < np; ++q) {
      const int i = P->order[q];
      if (P->nlev[i] > l) d.push_back(hier[i][l]);
    }
    V.np = (int)d.size();
    V.m = 0;
    while (V.m < V.np && P->nlev[P->order[V.m]] > l + 1) ++V.m;
    V.mc = V.m;
    while (V.mc < V.np && small(P->order[V.mc])) ++V.mc;
    if (l == 0 && V.mc > V.m) V.mc = V.m;   // the coarse kernel runs below level 0 only
    V.hg.resize(V.np);
    for (int q = 0; q < V.np; ++q) {
      PartGrid& g = V.hg[q];
      g.nx = d[q][0]; g.ny = d[q][1]; g.nz = d[q][2];
      const long long nn = (long long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
      const long long ne = (long long)g.nx * g.ny * g.nz;
      g.voff = V.n3; g.noff = V.nn; g.eoff = V.ne;
      g.uoff = P->uoff[P->order[q]];
      g.ueoff = P->ueoff[P->order[q]];
      V.n3 += align_up(3 * nn);
      V.nn += align_up(nn);
      V.ne += align_up(ne);
      if (q >= V.m && q < V.mc) {
        const long long work = std::max(ne, nn);
        V.coarse_threads = (int)std::max<long long>(
            V.coarse_threads, std::min<long long>(1024, (work + 63) / 64 * 64));
      }
    }
    std::vector<int> zc;
    part_zchunks(s->opt, d, zc);
    max_tiles = std::max<long long>(max_tiles, part_tiles(V, zc));
  }
  // grid transfers of the parts that go coarser
  for (int l = 0; l + 1 < L; ++l) {
    PLevel& V = P->lv[l];
    PLevel& C = P->lv[l + 1];
    V.hr.resize(V.m + 1);
    V.hp.resize(V.m + 1);
    int br = 0, bp = 0;
    // the level's work-group targets are shared among its parts by size (a
    // part of its own would get the whole target): one launch then dispatches
    // about as many work-groups as a single grid of the same total size
    double cn = 0.0, fn = 0.0;
    for (int q = 0; q < V.m; ++q) {
      const PartGrid& f = V.hg[q];
      const PartGrid& c = C.hg[q];
      cn += (double)(c.nx + 1) * (c.ny + 1) * (c.nz + 1);
      fn += (double)(f.nx + 1) * (f.ny + 1) * (f.nz + 1);
    }
    for (int q = 0; q < V.m; ++q) {
      const PartGrid& f = V.hg[q];
      const PartGrid& c = C.hg[q];
      PartTr& r = V.hr[q];
      r.Lcx = c.nx + 1; r.Lcy = c.ny + 1; r.Lcz = c.nz + 1;
      const double cshare = (double)r.Lcx * r.Lcy * r.Lcz / cn;
      r.g = tr_grid((long long)r.Lcx * r.Lcy, r.Lcz, false,
                    std::max(1, (int)(TR_TARGET_GROUPS * cshare + 0.5)));
      r.begin = br;
      r.src = f.voff; r.dst = c.voff; r.node = c.noff;
      br += r.g.gxy * r.g.nchunk;
      PartTr& p = V.hp[q];
      p.Lcx = r.Lcx; p.Lcy = r.Lcy; p.Lcz = r.Lcz;
      const double fshare = (double)(f.nx + 1) * (f.ny + 1) * (f.nz + 1) / fn;
      p.g = tr_grid(3LL * (f.nx + 1) * (f.ny + 1), f.nz + 1, true,
                    std::max(1, (int)(PROLONG_TARGET_GROUPS * fshare + 0.5)));
      p.begin = bp;
      p.src = c.voff; p.dst = f.voff; p.node = f.noff;
      bp += p.g.gxy * p.g.nchunk;
    }
    V.hr[V.m] = PartTr{};
    V.hr[V.m].begin = br;
    V.hp[V.m] = PartTr{};
    V.hp[V.m].begin = bp;
  }
  // device tables and buffers
  for (int l = 0; l < L; ++l) {
    PLevel& V = P->lv[l];
    CHK(upload(s, &V.dg, V.hg));
    CHK(upload(s, &V.dt, V.ht));
    if (l + 1 < L) {
      CHK(upload(s, &V.dr, V.hr));
      CHK(upload(s, &V.dp, V.hp));
    }
    CHK(dalloc(s, &V.mask, V.nn));
    CHK(dalloc(s, &V.mask_fixed, V.nn));
    CHK(dalloc(s, &V.mi64, V.nn));
    CHK(dalloc(s, &V.mi32, V.nn));
    CHK(dalloc(s, &V.E64, V.ne));
    CHK(dalloc(s, &V.E32, V.ne));
    if (f32) {
      if (l > 0) CHK(dalloc(s, &V.b32, V.n3));
      CHK(dalloc(s, &V.x32, V.n3));
      CHK(dalloc(s, &V.t32, V.n3));
      CHK(dalloc(s, &V.w32, V.n3));
    } else {
      if (l > 0) CHK(dalloc(s, &V.b64, V.n3));
      CHK(dalloc(s, &V.x64, V.n3));
      CHK(dalloc(s, &V.t64, V.n3));
      CHK(dalloc(s, &V.w64, V.n3));
    }
  }
  const size_t n0 = (size_t)P->lv[0].n3;
  CHK(dalloc(s, &P->u64, n0));
  CHK(dalloc(s, &P->r64, n0));
  CHK(dalloc(s, &P->fm64, n0));
  if (f32) {
    CHK(dalloc(s, &P->r32, n0));
    CHK(dalloc(s, &P->z32, n0));
    CHK(dalloc(s, &P->p32, n0));
    CHK(dalloc(s, &P->Ap32, n0));
    CHK(dalloc(s, &P->d32, n0));
  } else {
    CHK(dalloc(s, &P->z64, n0));
    CHK(dalloc(s, &P->p64, n0));
    CHK(dalloc(s, &P->Ap64, n0));
  }
  s->n_partial = (int)std::max<long long>(max_tiles, (long long)EW_BLOCKS * np);
  CHK(dalloc(s, &s->partial, (size_t)s->n_partial));
  CHK(dalloc(s, &s->dscal, (size_t)np));
  CHK(dalloc(s, &s->dflags, (size_t)np));
  if (hipHostMalloc((void**)&s->hscal, sizeof(Scalars) * np) != hipSuccess)
    return fail(s, VX_ERR_NOMEM, "pinned scalars");
  if (hipHostMalloc((void**)&s->hflags, sizeof(int) * np) != hipSuccess)
    return fail(s, VX_ERR_NOMEM, "pinned flags");
  HIPCHK(hipStreamSynchronize(s->stream));
  return VX_OK;
}

}  // namespace

extern "C" {

void vx_parts_destroy(vx_parts_handle P) {
  if (!P) return;
  vx_solver* s = &P->s;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  else (void)hipDeviceSynchronize();
  drop_graphs(s);
  if (s->cap_stream) (void)hipStreamDestroy(s->cap_stream);
  s->cap_stream = nullptr;
  for (void* q : s->allocs) (void)hipFree(q);
  s->allocs.clear();
  if (s->hscal) (void)hipHostFree(s->hscal);
  if (s->hflags) (void)hipHostFree(s->hflags);
  s->hscal = nullptr;
  s->hflags = nullptr;
  delete P;
}

int vx_parts_create(int32_t nparts, const int32_t* dims, double nu, const vx_options* opt,
                    vx_parts_handle* out) {
  if (!out) return VX_ERR_ARG;
  *out = nullptr;
  if (nparts < 1 || !dims || !(nu > -1.0 && nu < 0.5)) return VX_ERR_ARG;
  vx_options o;
  if (check_options(opt, 0, &o) != VX_OK) return VX_ERR_ARG;
  if (o.coarse_mode != VX_COARSE_SWEEPS || o.flexible) return VX_ERR_ARG;
  for (int i = 0; i < nparts; ++i) {
    if (dims[3 * i] < 1 || dims[3 * i + 1] < 1 || dims[3 * i + 2] < 1) return VX_ERR_ARG;
    const auto h = level_dims_rule(dims[3 * i], dims[3 * i + 1], dims[3 * i + 2], o);
    vx_options oi;
    if (check_options(&o, (int)h.size(), &oi) != VX_OK) return VX_ERR_ARG;
    // a one-level part smooths level 0 as its coarsest: its last step must
    // be a tile step, which carries the dot products of the PCG
    if (h.size() == 1 && o.coarse_sweeps < 2) return VX_ERR_ARG;
  }
  if (vx_device_count() < 1) return VX_ERR_NO_DEVICE;
  vx_parts* P = new vx_parts();
  vx_solver* s = &P->s;
  s->opt = o;
  s->opt.max_batch = nparts;
  s->nu = nu;
  s->K64 = elem_const(nu);
  s->K32 = {(float)s->K64.k1, (float)s->K64.k2, (float)s->K64.b1, (float)s->K64.b2,
            (float)s->K64.t1};
  s->kd = elem_diag(nu);
  s->lam_max = elem_lambda_max(nu) / s->kd;
  P->np = nparts;
  P->dims.resize(nparts);
  P->nlev.resize(nparts);
  for (int i = 0; i < nparts; ++i) P->dims[i] = {dims[3 * i], dims[3 * i + 1], dims[3 * i + 2]};
  const int rc = parts_build(P);
  if (rc != VX_OK) {
    vx_parts_destroy(P);
    return rc;
  }
  *out = P;
  return VX_OK;
}

const char* vx_parts_last_error(vx_parts_handle P) { return P ? P->s.err.c_str() : "null handle"; }

int vx_parts_set_stream(vx_parts_handle P, void* stream) {
  if (!P) return VX_ERR_ARG;
  P->s.stream = (hipStream_t)stream;
  return VX_OK;
}

int vx_parts_num_levels(vx_parts_handle P, int32_t part) {
  if (!P || part < 0 || part >= P->np) return VX_ERR_ARG;
  return P->nlev[part];
}

int vx_parts_set_fixed(vx_parts_handle P, const int64_t* fixed, const int64_t* counts) {
  if (!P || !counts) return VX_ERR_ARG;
  vx_solver* s = &P->s;
  PLevel& V0 = P->lv[0];
  std::vector<long long> start(P->np + 1, 0);
  for (int i = 0; i < P->np; ++i) {
    if (counts[i] < 0) return fail(s, VX_ERR_ARG, "negative fixed-DOF count");
    start[i + 1] = start[i] + counts[i];
  }
  if (start[P->np] > 0 && !fixed) return fail(s, VX_ERR_ARG, "null fixed DOFs");
  std::vector<uint8_t> m((size_t)V0.nn, 0);
  for (int q = 0; q < V0.np; ++q) {
    const int i = P->order[q];
    const PartGrid& g = V0.hg[q];
    const long long nn = (long long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
    std::fill(m.begin() + g.noff, m.begin() + g.noff + nn, (uint8_t)7);
    for (long long k = start[i]; k < start[i + 1]; ++k) {
      const int64_t d = fixed[k];
      if (d < 0 || d >= 3 * nn)
        return fail(s, VX_ERR_ARG, "fixed DOF id out of range on part " + std::to_string(i));
      m[g.noff + d / 3] &= (uint8_t)~(1u << (d % 3));
    }
  }
  HIPCHK(hipMemcpyAsync(V0.mask_fixed, m.data(), m.size(), hipMemcpyHostToDevice, s->stream));
  // injection of the fixed mask to the coarser levels (core/multigrid.py:255-278)
  for (size_t l = 1; l < P->lv.size(); ++l) {
    PLevel& F = P->lv[l - 1];
    PLevel& C = P->lv[l];
    for (int q = 0; q < C.np; ++q) {
      const PartGrid& c = C.hg[q];
      const long long nc = (long long)(c.nx + 1) * (c.ny + 1) * (c.nz + 1);
      hipLaunchKernelGGL(k_inject_mask, dim3(ew_grid(nc)), dim3(EW_THREADS), 0, s->stream,
                         c.nx + 1, c.ny + 1, c.nz + 1, F.mask_fixed + F.hg[q].noff,
                         C.mask_fixed + c.noff);
      LAUNCHCHK();
    }
  }
  for (auto& V : P->lv)
    HIPCHK(hipMemcpyAsync(V.mask, V.mask_fixed, V.nn, hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  P->fixed_set = true;
  P->density_set = false;
  return VX_OK;
}

int vx_parts_set_density(vx_parts_handle P, const double* E, const uint8_t* active, double e_min) {
  if (!P || !E) return VX_ERR_ARG;
  vx_solver* s = &P->s;
  if (!P->fixed_set) return fail(s, VX_ERR_ARG, "call vx_parts_set_fixed() before the density");
  const bool active_sys = e_min < 0.0;
  PLevel& V0 = P->lv[0];
  for (int q = 0; q < V0.np; ++q) {
    const PartGrid& g = V0.hg[q];
    const long long ne = (long long)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_import_E, dim3(ew_grid(ne)), dim3(EW_THREADS), 0, s->stream, g.nx, g.ny,
                       g.nz, E + g.ueoff, active ? active + g.ueoff : nullptr,
                       active_sys ? 0.0 : e_min, V0.E64 + g.eoff, V0.E32 + g.eoff);
    LAUNCHCHK();
  }
  for (size_t l = 1; l < P->lv.size(); ++l) {
    PLevel& F = P->lv[l - 1];
    PLevel& C = P->lv[l];
    for (int q = 0; q < C.np; ++q) {
      const PartGrid& c = C.hg[q];
      hipLaunchKernelGGL(k_coarsen_E<double>, dim3(ew_grid((long long)c.nx * c.ny * c.nz)),
                         dim3(EW_THREADS), 0, s->stream, c.nx, c.ny, c.nz, F.E64 + F.hg[q].eoff,
                         C.E64 + c.eoff, s->opt.e_coarsen);
      LAUNCHCHK();
    }
    hipLaunchKernelGGL((k_convert<double, float>), dim3(ew_grid(C.ne)), dim3(EW_THREADS), 0,
                       s->stream, C.ne, C.E64, C.E32);
    LAUNCHCHK();
  }
  for (auto& V : P->lv)
    for (int q = 0; q < V.np; ++q) {
      const PartGrid& g = V.hg[q];
      hipLaunchKernelGGL(k_effective_mask,
                         dim3(ew_grid((long long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1))),
                         dim3(EW_THREADS), 0, s->stream, g.nx, g.ny, g.nz, V.E64 + g.eoff,
                         V.mask_fixed + g.noff, active_sys ? 1 : 0, V.mask + g.noff);
      const dim3 g3((g.nx + 1 + 63) / 64, g.ny + 1, g.nz + 1);
      hipLaunchKernelGGL(k_node_minv<double>, g3, dim3(64), 0, s->stream, g.nx, g.ny, g.nz,
                         V.E64 + g.eoff, s->kd, V.mi64 + g.noff);
      hipLaunchKernelGGL(k_node_minv<float>, g3, dim3(64), 0, s->stream, g.nx, g.ny, g.nz,
                         V.E32 + g.eoff, (float)s->kd, V.mi32 + g.noff);
      LAUNCHCHK();
    }
  P->density_set = true;
  return VX_OK;
}

// Tile z-chunks of the parts, level by level.  Each candidate z-chunk (applied
// to every part of the level, at most the part's node planes; the rule's
// choice competes too) is timed on a smoothing step over all the parts of
// the level; the first within TUNE_TOL of the fastest wins (the rule's
// choice first, then ascending), so near-ties resolve the same way each run.
int vx_parts_autotune(vx_parts_handle P, int32_t reps) {
  if (!P) return VX_ERR_ARG;
  vx_solver* s = &P->s;
  if (!P->fixed_set || !P->density_set)
    return fail(s, VX_ERR_ARG, "vx_parts_autotune before vx_parts_set_fixed / set_density");
  if (s->opt.z_chunk > 0) return VX_OK;   // fixed by the options
  if (reps < 1) reps = 5;
  drop_graphs(s);   // the launch shapes change
  const bool f32 = s->opt.precision == VX_MIXED;
  const int cands[] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32};
  // candidate z-chunk sets per level, and a partial buffer for the largest
  // work-group count among them (the dot-product variants write one each)
  std::vector<std::vector<std::vector<int>>> cs(P->lv.size());
  long long need = s->n_partial;
  for (size_t l = 0; l < P->lv.size(); ++l) {
    PLevel& V = P->lv[l];
    std::vector<int> rule(V.np);
    int maxz = 0;
    for (int q = 0; q < V.np; ++q) {
      rule[q] = V.ht[q].zc;
      maxz = std::max(maxz, V.hg[q].nz + 1);
    }
    cs[l].push_back(rule);
    for (int c : cands) {
      if (c >= maxz) break;
      std::vector<int> z(V.np);
      for (int q = 0; q < V.np; ++q) z[q] = std::min(c, V.hg[q].nz + 1);
      if (std::find(cs[l].begin(), cs[l].end(), z) == cs[l].end()) cs[l].push_back(z);
    }
    std::vector<int> keep = rule;
    for (const auto& z : cs[l]) need = std::max<long long>(need, part_tiles(V, z));
    part_tiles(V, keep);
  }
  if (need > s->n_partial) {
    s->partial = nullptr;   // the smaller buffer stays owned by s->allocs
    CHK(dalloc(s, &s->partial, (size_t)need));
    s->n_partial = (int)need;
  }
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  int rc = VX_OK;
  auto put = [&](PLevel& V) -> int {
    HIPCHK(hipMemcpyAsync(V.dt, V.ht.data(), sizeof(PartTiles) * V.ht.size(),
                          hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));   // V.ht is rewritten next
    return VX_OK;
  };
  for (size_t l = 0; l < P->lv.size() && rc == VX_OK; ++l) {
    PLevel& V = P->lv[l];
    const PartsTarget t{P, (int)l, 0, V.np, nullptr};
    auto launch = [&]() -> int {
      SmoothCoef sc;
      sc.omega = s->opt.omega;
      sc.cu = 0.5;
      sc.cr = 0.1;
      if (f32) {
        sc.up = V.w32;
        return ttile<float, M_CHEB, false>(t, V.x32, l ? V.b32 : P->r32, V.t32, nullptr, &sc);
      }
      sc.up = V.w64;
      return ttile<double, M_CHEB, false>(t, V.x64, l ? V.b64 : P->r64, V.t64, nullptr, &sc);
    };
    std::vector<std::pair<size_t, float>> times;
    float best = INFINITY;
    for (size_t k = 0; k < cs[l].size() && rc == VX_OK; ++k) {
      part_tiles(V, cs[l][k]);
      rc = put(V);
      if (rc == VX_OK) rc = launch();   // warm-up
      if (rc != VX_OK) break;
      if (hipEventRecord(a, s->stream) != hipSuccess) { rc = fail(s, VX_ERR_HIP, "event"); break; }
      for (int r = 0; r < reps && rc == VX_OK; ++r) rc = launch();
      if (rc != VX_OK) break;
      float ms = 0.f;
      if (hipEventRecord(b, s->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
          hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        rc = fail(s, VX_ERR_HIP, "event timing");
        break;
      }
      times.push_back({k, ms});
      best = std::min(best, ms);
    }
    size_t pick = 0;
    for (const auto& kt : times)
      if (kt.second <= best * (1.0f + TUNE_TOL)) {
        pick = kt.first;
        break;
      }
    // the launch sizes come from the host table, the kernels' part lookup
    // from the device one: leave them equal, also after a failure
    part_tiles(V, cs[l][pick]);
    const int rp = put(V);
    if (rc == VX_OK) rc = rp;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

// tuned (or rule) z-chunk of one part's tile kernels at one level
int vx_parts_zchunk(vx_parts_handle P, int32_t part, int32_t level) {
  if (!P || part < 0 || part >= P->np || level < 0 || level >= P->nlev[part]) return VX_ERR_ARG;
  int q = 0;
  while (P->order[q] != part) ++q;
  return P->lv[level].ht[q].zc;
}

// Every part solved at once: F / U are the parts' full-grid float64 vectors
// packed one after another in caller order; each part follows exactly the
// recurrence of its own vx_solve (FP64) or iterative refinement (mixed) and
// stops at its own convergence.
int vx_solve_parts(vx_parts_handle P, const double* F, double* U, double tol, int32_t max_iter,
                   vx_progress_cb cb, void* user, double* hist_host, int32_t hist_cap,
                   vx_info* infos) {
  if (!P || !F || !U) return VX_ERR_ARG;
  vx_solver* s = &P->s;
  if (!P->fixed_set || !P->density_set)
    return fail(s, VX_ERR_ARG, "call vx_parts_set_fixed() and vx_parts_set_density() first");
  if (max_iter <= 0) max_iter = 1000;   // core/multigrid.py:481
  const int nb = P->np;
  s->nb = nb;
  const bool mixed = s->opt.precision == VX_MIXED;
  PLevel& V0 = P->lv[0];
  const long long n0 = V0.n3;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, s->stream));
  // f masked into this handle's packing; fnorm per part (core/multigrid.py:460-471)
  hipLaunchKernelGGL((k_mask_copy_parts<double, double>), dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0,
                     s->stream, V0.dg, 0, V0.mask, F, P->fm64, s->partial);
  LAUNCHCHK();
  CHK(pfinalize(P, false, F_FNORM));
  HIPCHK(hipMemsetAsync(P->u64, 0, sizeof(double) * n0, s->stream));
  CHK(sync_scalars(s));
  std::vector<double> fnorm(nb);
  std::vector<int> zero(nb, 0);
  for (int q = 0; q < nb; ++q) {
    fnorm[q] = s->hscal[q].fnorm;
    zero[q] = fnorm[q] == 0.0;
  }
  std::vector<double> hist;
  if (hist_host && hist_cap > 0) hist.assign((size_t)nb * hist_cap, 0.0);
  double* hp = hist.empty() ? nullptr : hist.data();
  std::vector<vx_info> inf(nb);
  for (int q = 0; q < nb; ++q) {
    inf[q] = vx_info{};
    inf[q].n_levels = P->nlev[P->order[q]];
    if (zero[q]) {   // zero load: u = 0 (core/multigrid.py:471-473)
      inf[q].converged = 1;
      inf[q].resid_ratio = inf[q].true_resid_ratio = 0.0;
    }
  }
  int rc = VX_OK;
  std::vector<PcgResult> pr(nb);
  std::vector<int> stalled(nb, 0);
  const PartsTarget all{P, 0, 0, nb, nullptr};
  if (!std::all_of(zero.begin(), zero.end(), [](int z) { return z; })) {
    if (!mixed) {
      rc = copy_vec<double>(s, P->r64, P->fm64, n0);
      std::vector<double> one(nb, 1.0);
      std::vector<int> off(nb, 0);
      if (rc == VX_OK)
        rc = ppcg<double>(P, P->u64, P->r64, P->z64, P->p64, P->Ap64, V0.t64, V0.w64,
                          std::vector<double>(nb, tol), max_iter, true, one, tol, off, cb, user, hp, hist_cap, pr, zero);
      for (int q = 0; q < nb; ++q) {
        if (zero[q]) continue;
        inf[q].iters = pr[q].iters;
        inf[q].converged = pr[q].converged ? 1 : 0;
        inf[q].resid_ratio = pr[q].ratio;
        stalled[q] = pr[q].stalled;
      }
    } else {
      // iterative refinement: FP64 true residual, FP32 inner MG-PCG per part
      std::vector<int> total(nb, 0), done(zero);
      std::vector<double> true_ratio(nb, 0.0);
      for (int outer = 0; outer <= s->opt.max_outer && rc == VX_OK; ++outer) {
        rc = ttile<double, M_RESID, true>(all, P->u64, P->fm64, P->r64, nullptr, nullptr);
        if (rc == VX_OK) rc = pfinalize(P, true, F_SET_TMP);
        if (rc == VX_OK) rc = sync_scalars(s);
        if (rc != VX_OK) break;
        int open = 0, most = 0;
        for (int q = 0; q < nb; ++q) {
          if (zero[q]) continue;
          true_ratio[q] = std::sqrt(s->hscal[q].tmp) / fnorm[q];
          inf[q].outer_iters = outer;
          if (true_ratio[q] < tol) done[q] = 1;
          if (!done[q]) ++open;
          most = std::max(most, total[q]);
        }
        if (!open || outer == s->opt.max_outer || most >= max_iter) break;
        for (int q = 0; q < nb; ++q) s->hscal[q].tmp2 = done[q] ? 1.0 : 0.0;
        HIPCHK(hipMemcpyAsync(s->dscal, s->hscal, sizeof(Scalars) * nb, hipMemcpyHostToDevice,
                              s->stream));
        hipLaunchKernelGGL(k_scale_to_f32_parts, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0,
                           s->stream, V0.dg, s->dscal, P->r64, P->r32);
        LAUNCHCHK();
        HIPCHK(hipMemsetAsync(P->d32, 0, sizeof(float) * n0, s->stream));
        hipLaunchKernelGGL(k_set_fnorm, dim3(nb), dim3(1), 0, s->stream, s->dscal, 1.0);
        LAUNCHCHK();
        // each part's inner solve stops where its own true residual would
        // reach half the tolerance, but not beyond inner_tol
        std::vector<double> want(nb, s->opt.inner_tol);
        for (int q = 0; q < nb; ++q)
          if (!done[q]) want[q] = std::max(s->opt.inner_tol, 0.5 * tol / true_ratio[q]);
        rc = ppcg<float>(P, P->d32, P->r32, P->z32, P->p32, P->Ap32, V0.t32, V0.w32, want,
                         max_iter - most, false, true_ratio, tol, total, cb, user, hp, hist_cap,
                         pr, done);
        if (rc != VX_OK) break;
        bool progressed = false;
        for (int q = 0; q < nb; ++q) {
          if (done[q]) continue;
          if (pr[q].iters > 0 && !std::isfinite(pr[q].ratio)) {
            rc = fail(s, VX_ERR_HIP, "non-finite residual in the FP32 inner solve");
            break;
          }
          total[q] += pr[q].iters;
          progressed = progressed || pr[q].iters > 0;
        }
        if (rc != VX_OK) break;
        hipLaunchKernelGGL(k_add_scaled_parts, dim3(EW_BLOCKS, nb), dim3(EW_THREADS), 0,
                           s->stream, V0.dg, s->dscal, P->d32, P->u64);
        LAUNCHCHK();
        if (!progressed) break;
      }
      for (int q = 0; q < nb; ++q) {
        if (zero[q]) continue;
        inf[q].iters = total[q];
        inf[q].resid_ratio = inf[q].true_resid_ratio = true_ratio[q];
        inf[q].converged = true_ratio[q] < tol ? 1 : 0;
      }
    }
  }
  // final mask of u (core/multigrid.py:547), the true FP64 residual, u out
  if (rc == VX_OK) {
    hipLaunchKernelGGL((k_mask_copy_parts<double, double>), dim3(EW_BLOCKS, nb), dim3(EW_THREADS),
                       0, s->stream, V0.dg, 2, V0.mask, P->u64, P->u64, nullptr);
    hipLaunchKernelGGL((k_mask_copy_parts<double, double>), dim3(EW_BLOCKS, nb), dim3(EW_THREADS),
                       0, s->stream, V0.dg, 1, V0.mask, P->u64, U, nullptr);
    LAUNCHCHK();
    HIPCHK(hipEventRecord(e1, s->stream));
    if (!mixed) {
      rc = ttile<double, M_RESID, true>(all, P->u64, P->fm64, P->z64, nullptr, nullptr);
      if (rc == VX_OK) rc = pfinalize(P, true, F_SET_TMP);
    }
    if (rc == VX_OK) rc = sync_scalars(s);
    if (rc == VX_OK) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      for (int q = 0; q < nb; ++q) {
        if (!mixed && !zero[q]) inf[q].true_resid_ratio = std::sqrt(s->hscal[q].tmp) / fnorm[q];
        inf[q].solve_ms = ms;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // back to the caller's part order
  for (int q = 0; q < nb; ++q) {
    const int i = P->order[q];
    if (infos) infos[i] = inf[q];
    if (hp)
      std::copy(hist.begin() + (size_t)q * hist_cap, hist.begin() + (size_t)(q + 1) * hist_cap,
                hist_host + (size_t)i * hist_cap);
  }
  s->nb = 1;
  if (rc != VX_OK) return rc;
  if (std::any_of(stalled.begin(), stalled.end(), [](int v) { return v; })) return VX_ERR_STALL;
  for (int q = 0; q < nb; ++q)
    if (!inf[q].converged) return VX_ERR_NOT_CONVERGED;
  return VX_OK;
}

}  // extern "C"
