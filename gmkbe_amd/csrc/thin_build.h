// Family construction for one tuple (see thin.h).  The census of miss types
// (how many vertices of side B have each type, and the lists) is taken either
// sequentially by one lane (build_families, also the CPU model) or by a whole
// wave (thin_kernels.hip); emit_families then enumerates the maximal count
// vectors and writes one family per vector.
#pragma once
#include "thin.h"

namespace gmk {
namespace thin {

// Upper bound on min |free \ N(w)| over w in side A outside L, from a bounded
// scan: the candidates are the neighbours of the free member of least degree
// (among the first 16), and each candidate's misses are counted by binary
// search until they reach the best seen.  Returns k + 1 when no w with at most
// k misses was seen.  Any value returned is attained by a real w, so a family
// found dominated with it is non-maximal (a witness exists); the bound only
// decides how many families are recognised.  cand0/cstep split the candidates
// between cooperating lanes (0/1 for one lane).
GMK_HD u32 dominance_bound(const Graph& g, const u32* L, u32 j, u32 k, const u32* freel, u32 nfree, u32 cand0,
                           u32 cstep, Cost& cost) {
  if (nfree == 0) return (g.a_hi - g.a_lo) > j ? 0u : k + 1;
  u32 vmin = freel[0], dmin = g.off[vmin + 1] - g.off[vmin];
  for (u32 i = 1; i < nfree && i < 16; ++i) {
    const u32 v = freel[i], dv = g.off[v + 1] - g.off[v];
    if (dv < dmin) { vmin = v; dmin = dv; }
  }
  cost.bytes += 8ull * (nfree < 16 ? nfree : 16);
  const u32* nv = g.adj + g.off[vmin];
  u32 best = k + 1;
  u32 probes = 0;  // work bound: binary searches per lane
  for (u32 e = cand0; e < dmin && best > 0 && probes < 4096; e += cstep) {
    const u32 w = nv[e];
    cost.bytes += 4;
    if (in_tuple(L, j, w)) continue;
    const u32* nw = g.adj + g.off[w];
    const u32 dw = g.off[w + 1] - g.off[w];
    u32 miss = 0, i = 0;
    for (; i < nfree && miss < best && probes < 4096; ++i) {
      if (freel[i] == vmin) continue;
      ++probes;
      if (!contains(nw, dw, freel[i], cost)) ++miss;
    }
    if (i == nfree && miss < best) best = miss;
  }
  return best;
}

// Enumerates the maximal count vectors of a tuple and writes its families.
//   mcnt[m]  number of side-B vertices of type m (m < 2^j; mcnt[full] unused)
//   far      number of side-B vertices adjacent to no tuple member
//   poolp[m] address of the list of type-m vertices (m = 0 is the free list)
//   nlp, d   N(L) ascending (needed only when the far pool is feasible)
//   dom      a w outside L misses dom free vertices (dominance_bound)
// A family whose chosen vertices all miss fewer than k tuple members, and with
// dom + (chosen vertices) <= k, is dropped: that w misses at most k vertices of
// every R of the family and no vertex it misses is saturated, so w can be added
// and no member is maximal.
// JM bounds the tuple size (j <= JM) so the per-lane arrays have 2^JM entries.
template <int JM, class Fams>
GMK_HD void emit_families(const u32* L, u32 j, u32 k, u32 min_b, u32 tuple_idx, const u32* mcnt, u64 far,
                          const u64* poolp, const u32* nlp, u32 d, u32 dom, const u32* fpick, u32 nfpick, Fams& fo,
                          Cost& cost) {
  const u32 full = (1u << j) - 1u;
  const u32 nfree = mcnt[0];
  u8 tmask[(1 << JM)];
  u32 nt = 0;
  for (u32 m = 1; m < full; ++m)
    if (mcnt[m] && (u32)popc(m) <= k) tmask[nt++] = m;
  const bool far_ok = j <= k && far > 0;
  if (far_ok) tmask[nt++] = full;  // the far pool is the last type
  auto psize_of = [&](u32 t) -> u64 { return (far_ok && t == nt - 1) ? far : (u64)mcnt[tmask[t]]; };
  {
    u64 maxr = nfree;
    for (u32 t = 0; t < nt; ++t) { const u64 ps = psize_of(t); maxr += ps < k ? ps : k; }
    if (maxr < min_b) return;
  }
  auto fill = [&](Family* f, u64 rs) {
    for (u32 i = 0; i < j; ++i) f->L[i] = L[i];
    f->j = j; f->nfree = nfree; f->rsize = (u32)rs;
    f->freep = (const u32*)poolp[0]; f->nlp = nlp; f->nl_size = d; f->tuple = tuple_idx;
    for (u32 q = 0; q < nfpick; ++q) f->fpick[q] = fpick[q];
    f->nfpick = nfpick;
    f->gmode = GUARD_NONE;  // the guard build (guard_build, thin_guard) comes later
    f->glog = 0;
    f->gmin = 0;
    f->s1cap = 0;
    f->gtab = 0;
  };
  if (nt == 0) {
    if (nfree >= min_b && nfree > 0 && dom > k) {
      Family* f = fo.next();
      cost.bytes += sizeof(Family);
      if (f) { fill(f, nfree); f->ntypes = 0; f->total = 1; }
    }
    return;
  }
  // Depth-first over the types, c[t] bounded by the members' remaining budgets;
  // a leaf is kept when every type with spare vertices has a member with no
  // budget left.
  u8 c[(1 << JM)];
  u32 budget[JM];
  for (u32 i = 0; i < j; ++i) budget[i] = k;
  auto cap_of = [&](u32 t) -> u32 {
    u32 cp = k;
    for (u32 i = 0; i < j; ++i)
      if ((tmask[t] >> i) & 1u) cp = budget[i] < cp ? budget[i] : cp;
    const u64 ps = psize_of(t);
    return ps < cp ? (u32)ps : cp;
  };
  auto apply = [&](u32 t, u32 v, bool add) {
    for (u32 i = 0; i < j; ++i)
      if ((tmask[t] >> i) & 1u) budget[i] = add ? budget[i] - v : budget[i] + v;
  };
  int t = 0;
  c[0] = cap_of(0);
  apply(0, c[0], true);
  for (;;) {
    if (t == (int)nt - 1) {
      bool maximal = true;
      for (u32 q = 0; q < nt && maximal; ++q) {
        if (c[q] < psize_of(q)) {
          bool blocked = false;
          for (u32 i = 0; i < j; ++i)
            if (((tmask[q] >> i) & 1u) && budget[i] == 0) blocked = true;
          if (!blocked) maximal = false;
        }
      }
      if (maximal) {
        u64 rs = nfree, total = 1;
        u32 used = 0;
        bool unsat = true;  // every chosen vertex misses fewer than k members
        for (u32 q = 0; q < nt; ++q) {
          rs += c[q];
          if (c[q]) {
            total = mul_sat(total, binom(psize_of(q), c[q]));
            ++used;
            if ((u32)popc(tmask[q]) >= k) unsat = false;
          }
        }
        const bool dominated = unsat && (u64)dom + (rs - nfree) <= k;
        if (rs >= min_b && rs > 0 && used <= (u32)MAXT && !dominated) {
          Family* f = fo.next();
          cost.bytes += sizeof(Family);
          if (f) {
            fill(f, rs);
            u32 n2 = 0;
            for (u32 q = 0; q < nt; ++q)
              if (c[q]) {
                f->mask[n2] = tmask[q];
                f->cnt[n2] = c[q];
                f->psize[n2] = (u32)psize_of(q);
                f->pool[n2] = (far_ok && q == nt - 1) ? FARP : poolp[tmask[q]];
                ++n2;
              }
            f->ntypes = n2;
            f->total = total;
          }
        }
      }
      for (;;) {
        if (t < 0) return;
        if (c[t] > 0) { apply(t, 1, false); --c[t]; break; }
        --t;
      }
      while (t < (int)nt - 1) { ++t; c[t] = cap_of(t); apply(t, c[t], true); }
    } else {
      ++t;
      c[t] = cap_of(t);
      apply(t, c[t], true);
    }
  }
}

// Sequential census by a j-way merge of the members' lists.  Returns false
// when the arena could not hold the tuple's lists (the caller retries with a
// smaller batch); a tuple that fails emits no family.
template <int JM, class Arena, class Fams>
GMK_HD bool build_families_t(const Graph& g, const u32* L, u32 j, u32 k, u32 min_b, u32 tuple_idx, Arena& ar,
                             Fams& fo, Cost& cost) {
  const u32 full = (1u << j) - 1u;
  u32 mcnt[1 << JM];
  u64 poolp[1 << JM];
  if (j == 1) {
    // one vertex: free list = N(s), the far pool is everything else
    const u32 s = L[0];
    const u32* ns = g.adj + g.off[s];
    const u32 d = g.off[s + 1] - g.off[s];
    cost.bytes += 8;
    mcnt[0] = d;
    poolp[0] = (u64)ns;
    const u32 dom = dominance_bound(g, L, 1, k, ns, d, 0, 1, cost);
    u32 fpick[4];
    const u32 nfp = free_picks(g.off, ns, d, k, fpick, cost);
    emit_families<JM>(L, 1, k, min_b, tuple_idx, mcnt, (u64)(g.b_hi - g.b_lo) - d, poolp, ns, d, dom, fpick, nfp,
                      fo, cost);
    return true;
  }
  u32 pos[JM], end[JM];
  for (u32 i = 0; i < j; ++i) { pos[i] = g.off[L[i]]; end[i] = g.off[L[i] + 1]; cost.bytes += 12 + 4ull * (end[i] - pos[i]); }
  for (u32 m = 0; m <= full; ++m) mcnt[m] = 0;
  u32 d = 0;
  for (;;) {
    u32 y = 0xFFFFFFFFu;
    for (u32 i = 0; i < j; ++i)
      if (pos[i] < end[i] && g.adj[pos[i]] < y) y = g.adj[pos[i]];
    if (y == 0xFFFFFFFFu) break;
    u32 hit = 0;
    for (u32 i = 0; i < j; ++i)
      if (pos[i] < end[i] && g.adj[pos[i]] == y) { hit |= 1u << i; ++pos[i]; }
    ++mcnt[full & ~hit];
    ++d;
  }
  const bool far_ok = j <= k;
  const u64 nl_off = far_ok ? ar.alloc(d) : 0;
  if (nl_off == SAT) return false;
  const u64 pool_off = ar.alloc(d);
  if (pool_off == SAT) return false;
  u32 moff[1 << JM];
  {
    u32 acc = 0;
    for (u32 m = 0; m <= full; ++m) { moff[m] = acc; acc += mcnt[m]; }
  }
  u32* nl = far_ok ? ar.ptr(nl_off) : nullptr;
  u32* pools = ar.ptr(pool_off);
  for (u32 m = 0; m <= full; ++m) poolp[m] = (u64)(pools + moff[m]);
  for (u32 i = 0; i < j; ++i) cost.bytes += 4ull * (end[i] - g.off[L[i]]);
  cost.bytes += (far_ok ? 8ull : 4ull) * d;
  for (u32 i = 0; i < j; ++i) pos[i] = g.off[L[i]];
  {
    u32 fillc[1 << JM];
    for (u32 m = 0; m <= full; ++m) fillc[m] = 0;
    u32 w = 0;
    for (;;) {
      u32 y = 0xFFFFFFFFu;
      for (u32 i = 0; i < j; ++i)
        if (pos[i] < end[i] && g.adj[pos[i]] < y) y = g.adj[pos[i]];
      if (y == 0xFFFFFFFFu) break;
      u32 hit = 0;
      for (u32 i = 0; i < j; ++i)
        if (pos[i] < end[i] && g.adj[pos[i]] == y) { hit |= 1u << i; ++pos[i]; }
      const u32 m = full & ~hit;
      if (nl) nl[w] = y;
      ++w;
      pools[moff[m] + fillc[m]++] = y;
    }
  }
  const u32 dom = dominance_bound(g, L, j, k, (const u32*)poolp[0], mcnt[0], 0, 1, cost);
  u32 fpick[4];
  const u32 nfp = free_picks(g.off, (const u32*)poolp[0], mcnt[0], k, fpick, cost);
  emit_families<JM>(L, j, k, min_b, tuple_idx, mcnt, (u64)(g.b_hi - g.b_lo) - d, poolp, nl, d, dom, fpick, nfp, fo,
                    cost);
  return true;
}

// Dispatch on the tuple size so the arrays are sized 2^j.
template <class Arena, class Fams>
GMK_HD bool build_families(const Graph& g, const u32* L, u32 j, u32 k, u32 min_b, u32 tuple_idx, Arena& ar,
                           Fams& fo, Cost& cost) {
  switch (j) {
    case 1: return build_families_t<1>(g, L, j, k, min_b, tuple_idx, ar, fo, cost);
    case 2: return build_families_t<2>(g, L, j, k, min_b, tuple_idx, ar, fo, cost);
    case 3: return build_families_t<3>(g, L, j, k, min_b, tuple_idx, ar, fo, cost);
    case 4: return build_families_t<4>(g, L, j, k, min_b, tuple_idx, ar, fo, cost);
    case 5: return build_families_t<5>(g, L, j, k, min_b, tuple_idx, ar, fo, cost);
    default: return build_families_t<MAXJ>(g, L, j, k, min_b, tuple_idx, ar, fo, cost);
  }
}

}  // namespace thin
}  // namespace gmk
