// Family construction for one tuple (see thin.h).  Shared by the device kernel
// (one lane per tuple) and the host-side model used by the tests; the arena and
// the family sink are supplied by the caller.
#pragma once
#include "thin.h"

namespace gmk {
namespace thin {

// Returns false when the arena could not hold the tuple's lists (the caller
// retries with a smaller batch); families of a tuple are emitted only after the
// arena allocation succeeded, so a failed tuple emits nothing.
template <class Arena, class Fams>
GMK_HD bool build_families(const Graph& g, const u32* L, u32 j, u32 k, u32 min_b, u32 tuple_idx, Arena& ar,
                           Fams& fo, Cost& cost) {
  const u32 full = (1u << j) - 1u;
  u32 pos[MAXJ], end[MAXJ];
  for (u32 i = 0; i < j; ++i) { pos[i] = g.off[L[i]]; end[i] = g.off[L[i] + 1]; cost.bytes += 12 + 4ull * (end[i] - pos[i]); }
  u32 mcnt[1 << MAXJ];
  for (u32 m = 0; m <= full; ++m) mcnt[m] = 0;
  u32 d = 0;
  // pass 1: sizes of N(L) and of every miss type
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
  const u64 far = (u64)(g.b_hi - g.b_lo) - d;
  // feasible non-free types and the free count
  const u32 nfree = mcnt[0];
  u32 tmask[(1 << MAXJ)];
  u32 nt = 0;
  for (u32 m = 1; m < full; ++m)
    if (mcnt[m] && (u32)popc(m) <= k) tmask[nt++] = m;
  const bool far_ok = j <= k && far > 0;
  // tmask[nt] is the far pool when far_ok (mask = full)
  if (far_ok) tmask[nt++] = full;
  // Quick bound: |R| can never reach min_b.
  {
    u64 maxr = nfree;
    for (u32 t = 0; t < nt; ++t) maxr += (tmask[t] == full && far_ok) ? (far < k ? far : k) : mcnt[tmask[t]];
    if (maxr < min_b) return true;
  }
  const u64 nl_off = ar.alloc(d);
  if (nl_off == SAT) return false;
  const u64 pool_off = ar.alloc(d);
  if (pool_off == SAT) return false;
  u32 moff[1 << MAXJ];
  {
    u32 acc = 0;
    for (u32 m = 0; m <= full; ++m) { moff[m] = acc; acc += mcnt[m]; }
  }
  u32* nl = ar.ptr(nl_off);
  u32* pools = ar.ptr(pool_off);
  // second merge pass reads the lists again and writes N(L) and the pools
  for (u32 i = 0; i < j; ++i) cost.bytes += 4ull * (end[i] - g.off[L[i]]);
  cost.bytes += 8ull * d;
  // pass 2: write N(L) (ascending) and the pools grouped by type (each ascending)
  for (u32 i = 0; i < j; ++i) pos[i] = g.off[L[i]];
  {
    u32 fill[1 << MAXJ];
    for (u32 m = 0; m <= full; ++m) fill[m] = 0;
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
      nl[w++] = y;
      pools[moff[m] + fill[m]++] = y;
    }
  }
  // Maximal count vectors: depth-first over the types, c[t] bounded by the
  // members' remaining budgets; a leaf is kept when every type with spare
  // vertices has a member with no budget left.
  u32 c[(1 << MAXJ)];
  u32 budget[MAXJ];
  for (u32 i = 0; i < j; ++i) budget[i] = k;
  auto psize_of = [&](u32 t) -> u64 { return (far_ok && t == nt - 1) ? far : (u64)mcnt[tmask[t]]; };
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
  if (nt == 0) {
    // only the free vertices: R = free
    if (nfree >= min_b && nfree > 0) {
      Family* f = fo.next();
      if (f) {
        for (u32 i = 0; i < j; ++i) f->L[i] = L[i];
        f->j = j; f->nfree = nfree; f->ntypes = 0; f->rsize = nfree;
        f->free_off = pool_off + moff[0]; f->nl_off = nl_off; f->nl_size = d; f->tuple = tuple_idx;
        f->total = 1;
      }
    }
    return true;
  }
  c[0] = cap_of(0);
  apply(0, c[0], true);
  for (;;) {
    if (t == (int)nt - 1) {
      // leaf: test maximality
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
        for (u32 q = 0; q < nt; ++q) {
          rs += c[q];
          if (c[q]) { total = mul_sat(total, binom(psize_of(q), c[q])); ++used; }
        }
        if (rs >= min_b && rs > 0 && used <= (u32)MAXT) {
          Family* f = fo.next();
          cost.bytes += sizeof(Family);
          if (f) {
            for (u32 i = 0; i < j; ++i) f->L[i] = L[i];
            f->j = j; f->nfree = nfree; f->rsize = (u32)rs;
            f->free_off = pool_off + moff[0]; f->nl_off = nl_off; f->nl_size = d; f->tuple = tuple_idx;
            u32 n2 = 0;
            for (u32 q = 0; q < nt; ++q)
              if (c[q]) {
                f->mask[n2] = tmask[q];
                f->cnt[n2] = c[q];
                f->psize[n2] = (u32)psize_of(q);
                f->poff[n2] = (far_ok && q == nt - 1) ? (u64)FAR : pool_off + moff[tmask[q]];
                ++n2;
              }
            f->ntypes = n2;
            f->total = total;
          }
        }
      }
      // backtrack: lower the deepest counter that can still be lowered
      for (;;) {
        if (t < 0) return true;
        if (c[t] > 0) {
          apply(t, 1, false);
          --c[t];
          break;
        }
        --t;
      }
      // descend with maximal counts on the remaining types
      while (t < (int)nt - 1) {
        ++t;
        c[t] = cap_of(t);
        apply(t, c[t], true);
      }
    } else {
      ++t;
      c[t] = cap_of(t);
      apply(t, c[t], true);
    }
  }
}

}  // namespace thin
}  // namespace gmk
