// Collision detection of one world: nxn and sweep-and-prune broadphase,
// primitive narrowphase and contact-parameter mixing.
//
// Each world's broadphase pairs and contacts occupy one contiguous block of
// the global pools (collision_pair / contact_*), reserved with one atomicAdd
// per wavefront on ncollision / ncon.  Inside a block the order is the
// world's pair enumeration order, so every world's contact set is
// deterministic; only the placement of the blocks depends on scheduling,
// as the reference's per-pair atomics do (collision_driver.py:54,
// collision_primitive.py:64).
#pragma once

#include "common.hpp"
#include "linalg_reg.hpp"

namespace mjh {

// collision_driver.py:27-49 _geom_filter
MJH_DEV bool geom_filter(const mjhModel& m, int g1, int g2, bool filterparent) {
  const int b1 = m.geom_bodyid[g1], b2 = m.geom_bodyid[g2];
  const int ct1 = m.geom_contype[g1], ct2 = m.geom_contype[g2];
  const int ca1 = m.geom_conaffinity[g1], ca2 = m.geom_conaffinity[g2];
  const int w1 = m.body_weldid[b1], w2 = m.body_weldid[b2];
  const int wp1 = m.body_weldid[m.body_parentid[w1]];
  const int wp2 = m.body_weldid[m.body_parentid[w2]];
  const bool self = w1 == w2;
  const bool parent_child = filterparent && w1 != 0 && w2 != 0 && (w1 == wp2 || w2 == wp1);
  const bool mask = (ct1 & ca2) || (ct2 & ca1);
  return mask && !self && !parent_child;
}

// upper-triangle pair index p -> (g1 < g2), lexicographic like
// collision_driver.py:465-478: a float estimate of the row, then exact
// integer corrections
MJH_DEV void pair_from_index(int p, int n, int* g1, int* g2) {
  // row r starts at r*(2n-r-1)/2
  const float disc = (float)(2 * n - 1) * (float)(2 * n - 1) - 8.0f * (float)p;
  int r = (int)(((float)(2 * n - 1) - sqrtf(fmaxf(disc, 0.0f))) * 0.5f);
  if (r < 0) r = 0;
  while (r > 0 && (long)r * (2 * n - r - 1) / 2 > p) r--;
  while ((long)(r + 1) * (2 * n - r - 2) / 2 <= p) r++;
  *g1 = r;
  *g2 = p - r * (2 * n - r - 1) / 2 + r + 1;
}

// The nxn broadphase's per-geom inputs, staged in LDS once per world
// (kNxnWords words per geom): the filter's ints (contype, conaffinity,
// weld body, weld body of the weld parent) and the bounding test's floats
// (position, rbound, plane normal = third column of geom_xmat, margin).
constexpr int kNxnWords = 12;

// The filter's ints are only staged when `ints` is set (no pair table).
template <int TEAM>
MJH_DEV void nxn_stage_geoms(const mjhModel& m, const float* gxpos, const float* gxmat, float* lds, bool ints,
                             const Team<TEAM>& t) {
  int* gi = (int*)lds;
  float* gf = lds + 4 * m.ngeom;
  for (int g = t.lane; g < m.ngeom; g += TEAM) {
    if (ints) {
      const int wd = m.body_weldid[m.geom_bodyid[g]];
      gi[4 * g] = m.geom_contype[g];
      gi[4 * g + 1] = m.geom_conaffinity[g];
      gi[4 * g + 2] = wd;
      gi[4 * g + 3] = m.body_weldid[m.body_parentid[wd]];
    }
    const float* x = gxmat + 9 * g;
    gf[8 * g] = gxpos[3 * g];
    gf[8 * g + 1] = gxpos[3 * g + 1];
    gf[8 * g + 2] = gxpos[3 * g + 2];
    gf[8 * g + 3] = m.geom_rbound[g];
    gf[8 * g + 4] = x[2];
    gf[8 * g + 5] = x[5];
    gf[8 * g + 6] = x[8];
    gf[8 * g + 7] = m.geom_margin[g];
  }
}

// collision_driver.py:480-508 bounding test of the nxn broadphase, on the
// staged geom table
MJH_DEV bool nxn_bounds(const float* lds, int ngeom, int g1, int g2) {
  const float* gf = lds + 4 * ngeom;
  const float* f1 = gf + 8 * g1;
  const float* f2 = gf + 8 * g2;
  const float s1 = f1[3], s2 = f2[3];
  const float bound = s1 + s2 + fmaxf(f1[7], f2[7]);
  const V3 dif = V3{f2[0], f2[1], f2[2]} - V3{f1[0], f1[1], f1[2]};
  if (s1 != 0.0f && s2 != 0.0f) return dot(dif, dif) <= bound * bound;
  if (s1 == 0.0f) return dot(dif, V3{f1[4], f1[5], f1[6]}) <= bound;
  return dot(-dif, V3{f2[4], f2[5], f2[6]}) <= bound;
}

// collision_driver.py:27-49 filter and the bounding test, on the staged
// geom table
MJH_DEV bool nxn_accept(const float* lds, int ngeom, int g1, int g2, bool filterparent) {
  const int* gi = (const int*)lds;
  const int ct1 = gi[4 * g1], ca1 = gi[4 * g1 + 1], w1 = gi[4 * g1 + 2], wp1 = gi[4 * g1 + 3];
  const int ct2 = gi[4 * g2], ca2 = gi[4 * g2 + 1], w2 = gi[4 * g2 + 2], wp2 = gi[4 * g2 + 3];
  const bool self = w1 == w2;
  const bool parent_child = filterparent && w1 != 0 && w2 != 0 && (w1 == wp2 || w2 == wp1);
  const bool mask = (ct1 & ca2) || (ct2 & ca1);
  if (!mask || self || parent_child) return false;
  return nxn_bounds(lds, ngeom, g1, g2);
}

// Candidate p of the nxn broadphase: entry p of the model's nxn_pair table
// (model-only filter done) when the model has one, else upper-triangle
// pair p with the full filter.  Sets (g1, g2) and returns whether the pair
// passes; with `test` unset only sets (g1, g2).
MJH_DEV bool nxn_candidate(const mjhModel& m, const float* lds, int p, bool table, bool fp, bool test, int* g1,
                           int* g2) {
  if (table) {
    const int e = m.nxn_pair[p];
    *g1 = e & 0x7fff;
    *g2 = (e >> 15) & 0x7fff;
    if (!test) return false;
    if (fp && ((e >> 30) & 1)) return false;
    return nxn_bounds(lds, m.ngeom, *g1, *g2);
  }
  pair_from_index(p, m.ngeom, g1, g2);
  return test && nxn_accept(lds, m.ngeom, *g1, *g2, fp);
}

// ---- sweep-and-prune helpers (collision_driver.py:71-301) ----
MJH_DEV float4 encode_plane(V3 n, V3 p, float margin) {
  n = normalize(n);
  const float off = -dot(n, p + n * margin);
  const float s = fabsf(off);
  if (fabsf(off) < 1e-6f) return make_float4(n.x, n.y, n.z, -1.0f);
  if (off < 0.0f) return make_float4(s * n.x, s * n.y, s * n.z, -2.0f);
  return make_float4(s * n.x, s * n.y, s * n.z, -3.0f);
}
MJH_DEV float4 decode_plane(float4 e) {
  const float mag = sqrtf(e.x * e.x + e.y * e.y + e.z * e.z + e.w * e.w);
  const V3 n = normalize(V3{e.x, e.y, e.z});
  if (e.w == -1.0f) return make_float4(n.x, n.y, n.z, 0.0f);
  if (e.w == -2.0f) return make_float4(n.x, n.y, n.z, -mag);
  return make_float4(n.x, n.y, n.z, mag);
}
MJH_DEV bool sap_overlap(float4 a, float4 b) {
  if (a.w < 0.0f && b.w < 0.0f) return false;
  if (a.w < 0.0f || b.w < 0.0f) {
    if (b.w < 0.0f) {
      const float4 tmp = a;
      a = b;
      b = tmp;
    }
    a = decode_plane(a);
    const float dist = b.x * a.x + b.y * a.y + b.z * a.z + a.w;
    return dist <= b.w;
  }
  const V3 dl = V3{a.x - b.x, a.y - b.y, a.z - b.z};
  const float rs = a.w + b.w;
  return dot(dl, dl) <= rs * rs;
}

// ---------------------------------------------------------------------------
// Broadphase.  The candidate enumeration is shared by the count pass and the
// write pass; f(g1, g2) is called for every accepted pair in order.
template <int TEAM>
struct PairWriter {
  const mjhModel& m;
  const mjhData& d;
  int base, cap, k;
  MJH_DEV void put(int g1, int g2, int world) {
    const int idx = base + k++;
    if (idx < cap) {
      if (m.geom_type[g1] > m.geom_type[g2]) {
        const int tmp = g1;
        g1 = g2;
        g2 = tmp;
      }
      d.collision_pair[2 * idx] = g1;
      d.collision_pair[2 * idx + 1] = g2;
      d.collision_worldid[idx] = world;
    }
  }
};

// collision_driver.py:458-512 nxn_broadphase.  lds: kNxnWords * ngeom
// floats.  Every pair is tested once: bit i of `hits` keeps the result of
// the lane's i-th pair for the write pass (pairs past the 32nd are tested
// again there).
template <int TEAM>
MJH_DEV void nxn_broadphase(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  const int w = t.world, ng = m.ngeom;
  const int npair = ng * (ng - 1) / 2;
  const bool fp = !(m.opt.disableflags & mjhDSBL_FILTERPARENT);
  const float* gxpos = WPTR(d.geom_xpos, w, ng * 3);
  const float* gxmat = WPTR(d.geom_xmat, w, ng * 9);
  // candidates: the model's filtered pair table, or every pair of a model
  // too large for one
  const bool table = ng <= MJH_NXN_TABLE_MAX;
  const int ncand = table ? m.nnxn : npair;
  if (t.valid) nxn_stage_geoms(m, gxpos, gxmat, lds, !table, t);
  sync();
  // pass 1: test and count
  int count = 0;
  unsigned hits = 0;
  if (t.valid)
    for (int p = t.lane, i = 0; p < ncand; p += TEAM, i++) {
      int g1, g2;
      if (nxn_candidate(m, lds, p, table, fp, true, &g1, &g2)) {
        count++;
        if (i < 32) hits |= 1u << i;
      }
    }
  const int total = team_sum_int<TEAM>(count);
  const int base = team_reserve<TEAM>(d.ncollision, t.valid ? total : 0);
  if (!t.valid) return;
  const int cap = d.nconmax;
  if (t.lane == 0) {
    d.world_pair[2 * w] = base;
    d.world_pair[2 * w + 1] = total <= 0 ? 0 : max(0, min(total, cap - base));
    if (base + total > cap) atomicAdd(d.overflow + 0, min(total, base + total - cap));
  }
  // pass 2: write in pair order
  int k = 0;
  for (int p0 = 0, i = 0; p0 < ncand; p0 += TEAM, i++) {
    const int p = p0 + t.lane;
    int g1 = 0, g2 = 0;
    bool ok = false;
    if (p < ncand) {
      ok = nxn_candidate(m, lds, p, table, fp, i >= 32, &g1, &g2);
      if (i < 32) ok = ((hits >> i) & 1u) != 0;
    }
    int tot;
    const int off = team_exscan_bits<TEAM, 1>(ok ? 1 : 0, &tot);
    if (ok) {
      PairWriter<TEAM> pw{m, d, base, cap, k + off};
      pw.put(g1, g2, w);
    }
    k += tot;
  }
}

// LDS radix sort of n (key, value) pairs by key, stable (the reference
// sorts with wp.utils.radix_sort_pairs): four passes of 8-bit digits,
// least significant first, each a 256-bucket histogram (LDS atomics), a
// team scan of the buckets and a scatter of the keys in chunks of TEAM in
// input order, a key's place among equal digits of its chunk taken from
// eight ballots.  Keys are floats mapped to order-preserving unsigned
// integers (-0 as +0); key2 / val2 are scratch of n, hist of 256.  The
// result is in key / val (an even number of passes).
MJH_DEV unsigned float_key(float f) {
  unsigned u = __float_as_uint(f == 0.0f ? 0.0f : f);
  return u ^ ((u >> 31) ? 0xffffffffu : 0x80000000u);
}
MJH_DEV float key_float(unsigned u) { return __uint_as_float(u ^ ((u >> 31) ? 0x80000000u : 0xffffffffu)); }

template <int TEAM>
MJH_DEV void radix_sort_lds(unsigned* key, int* val, unsigned* key2, int* val2, int* hist, int n,
                            const Team<TEAM>& t) {
  const unsigned long long team = team_lanes<TEAM>();
  const unsigned long long below = __lanemask_lt() & team;
  for (int shift = 0; shift < 32; shift += 8) {
    for (int b = t.lane; b < 256; b += TEAM) hist[b] = 0;
    sync();
    if (t.valid)
      for (int i = t.lane; i < n; i += TEAM) atomicAdd(&hist[(key[i] >> shift) & 255u], 1);
    sync();
    {
      constexpr int per = 256 / TEAM;
      int local[per];
      int sum = 0;
#pragma unroll
      for (int q = 0; q < per; q++) {
        local[q] = hist[t.lane * per + q];
        sum += local[q];
      }
      int tot;
      int run = team_exscan_int<TEAM>(sum, t.lane, &tot);
#pragma unroll
      for (int q = 0; q < per; q++) {
        hist[t.lane * per + q] = run;
        run += local[q];
      }
    }
    sync();
    for (int i0 = 0; i0 < n; i0 += TEAM) {
      const int i = i0 + t.lane;
      const bool in = t.valid && i < n;
      const unsigned k = in ? key[i] : 0u;
      const int v = in ? val[i] : 0;
      const unsigned dg = (k >> shift) & 255u;
      unsigned long long peers = __ballot(in) & team;
#pragma unroll
      for (int b = 0; b < 8; b++) {
        const unsigned long long m = __ballot((dg >> b) & 1u);
        peers &= ((dg >> b) & 1u) ? m : ~m;
      }
      int dst = 0;
      if (in) dst = hist[dg] + __popcll(peers & below);
      wave_fence();
      // the highest lane of each digit's peers moves the bucket on
      if (in && (peers & ~(below | (1ull << (threadIdx.x & 63)))) == 0) hist[dg] += __popcll(peers);
      if (in) {
        key2[dst] = k;
        val2[dst] = v;
      }
      wave_fence();
    }
    sync();
    unsigned* tk = key;
    key = key2;
    key2 = tk;
    int* tv = val;
    val = val2;
    val2 = tv;
  }
}

// LDS words of sap_broadphase: the sort's keys, values, their scratch
// copies and 256 buckets, then (16-byte aligned) a float4 per geom for the
// sorted bounding volumes.
__host__ __device__ inline int sap_sort_words(int ng) { return (4 * ng + 256 + 3) & ~3; }
__host__ __device__ inline int lds_sap_words(int ng) { return sap_sort_words(ng) + 4 * ng; }

// collision_driver.py:350-455 sap_broadphase.  lds: lds_sap_words(ngeom).
template <int TEAM>
MJH_DEV void sap_broadphase(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  const int w = t.world, ng = m.ngeom;
  float* key = lds;
  int* val = reinterpret_cast<int*>(lds + ng);
  unsigned* ukey = reinterpret_cast<unsigned*>(key);
  unsigned* key2 = reinterpret_cast<unsigned*>(lds + 2 * ng);
  int* val2 = reinterpret_cast<int*>(lds + 3 * ng);
  int* hist = reinterpret_cast<int*>(lds + 4 * ng);
  float4* gs = reinterpret_cast<float4*>(lds + sap_sort_words(ng));
  const bool fp = !(m.opt.disableflags & mjhDSBL_FILTERPARENT);
  const float* gxpos = WPTR(d.geom_xpos, w, ng * 3);
  const float* gxmat = WPTR(d.geom_xmat, w, ng * 9);
  float* lower = WPTR(d.sap_projection_lower, w, ng);
  float* upper = WPTR(d.sap_projection_upper, w, ng);
  int* sidx = WPTR(d.sap_sort_index, w, ng);
  // fixed sweep direction normalize(0.5935, 0.7790, 0.1235)
  const V3 dir = normalize(V3{0.5935f, 0.7790f, 0.1235f});
  if (t.valid)
    for (int i = t.lane; i < ng; i += TEAM) {
      const V3 c = ld3(gxpos + 3 * i);
      float r = m.geom_rbound[i];
      if (r == 0.0f) r = 1000000000.0f;
      const float rad = r + m.geom_margin[i];
      const float center = dot(dir, c);
      ukey[i] = float_key(center - rad);
      val[i] = i;
      upper[i] = center + rad;
    }
  sync();
  radix_sort_lds(ukey, val, key2, val2, hist, ng, t);
  if (t.valid)
    for (int i = t.lane; i < ng; i += TEAM) key[i] = key_float(ukey[i]);
  sync();
  if (t.valid)
    for (int i = t.lane; i < ng; i += TEAM) {
      lower[i] = key[i];
      sidx[i] = val[i];
      const int g = val[i];
      const V3 c = ld3(gxpos + 3 * g);
      const float r = m.geom_rbound[g];
      const float mg = m.geom_margin[g];
      float4 s;
      if (r == 0.0f) {
        const float* x = gxmat + 9 * g;
        s = encode_plane(V3{x[2], x[5], x[8]}, c, mg);
      } else {
        s = make_float4(c.x, c.y, c.z, r + mg);
      }
      gs[i] = s;
      reinterpret_cast<float4*>(WPTR(d.sap_geom_sort, w, ng * 4))[i] = s;
    }
  sync();
  // candidate range per sorted geom: j in (i, limit] (collision_driver.py:220-237)
  int* range = WPTR(d.sap_range, w, ng);
  if (t.valid)
    for (int i = t.lane; i < ng; i += TEAM) {
      const float end = upper[val[i]];
      int lo = i + 1, hi = ng;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (key[mid] > end) hi = mid;
        else lo = mid + 1;
      }
      const int limit = min(ng - 1, lo);
      range[i] = limit - i;
    }
  sync();
  // per-world inclusive scan of the ranges (the reference scans all worlds),
  // in chunks of TEAM; kept in LDS over the sorted keys (no longer needed)
  int* cs = reinterpret_cast<int*>(key);
  {
    int run = 0;
    for (int i0 = 0; i0 < ng; i0 += TEAM) {
      const int i = i0 + t.lane;
      const int r = (t.valid && i < ng) ? range[i] : 0;
      int tot;
      const int ex = team_exscan_int<TEAM>(r, t.lane, &tot);
      if (t.valid && i < ng) {
        cs[i] = run + ex + r;
        d.sap_cumulative_sum[(size_t)w * ng + i] = run + ex + r;
      }
      run += tot;
    }
  }
  sync();
  // candidate pairs p in [0, P), P = cs[ng - 1], i-major: sorted geom i
  // (cs[i - 1] <= p < cs[i]) and j = i + 1 + p - cs[i - 1]; lanes take
  // pairs, so a plane whose range spans every geom costs no lane more than
  // the others
  const int P = (t.valid && ng > 0) ? cs[ng - 1] : 0;
  auto pair_of = [&](int p, int* pi, int* pj) {
    int lo = 0, hi = ng - 1;  // first i with cs[i] > p
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (cs[mid] > p) hi = mid;
      else lo = mid + 1;
    }
    *pi = lo;
    *pj = lo + 1 + p - (lo > 0 ? cs[lo - 1] : 0);
  };
  // pass 1: count accepted pairs
  int count = 0;
  if (t.valid)
    for (int p = t.lane; p < P; p += TEAM) {
      int i, j;
      pair_of(p, &i, &j);
      if (geom_filter(m, val[i], val[j], fp) && sap_overlap(gs[i], gs[j])) count++;
    }
  const int total = team_sum_int<TEAM>(count);
  const int base = team_reserve<TEAM>(d.ncollision, t.valid ? total : 0);
  const int cap = d.nconmax;
  if (t.valid && t.lane == 0) {
    d.world_pair[2 * w] = base;
    d.world_pair[2 * w + 1] = total <= 0 ? 0 : max(0, min(total, cap - base));
    if (base + total > cap) atomicAdd(d.overflow + 0, min(total, base + total - cap));
  }
  // pass 2: write, ordered by (sorted i, j) = by p
  int k = 0;
  for (int p0 = 0; p0 < P; p0 += TEAM) {
    const int p = p0 + t.lane;
    int i = 0, j = 0;
    bool acc = false;
    if (t.valid && p < P) {
      pair_of(p, &i, &j);
      acc = geom_filter(m, val[i], val[j], fp) && sap_overlap(gs[i], gs[j]);
    }
    int tot;
    const int off = team_exscan_bits<TEAM, 1>(acc ? 1 : 0, &tot);
    if (acc) {
      PairWriter<TEAM> pw{m, d, base, cap, k + off};
      pw.put(val[i], val[j], w);
    }
    k += tot;
  }
  sync();
}

// ---------------------------------------------------------------------------
// Narrowphase (collision_primitive.py:52-268).  A primitive emits up to four
// contacts; ContactSink either counts them or writes them.
struct ContactSink {
  const mjhData* d;
  int base, cap, k, world;
  bool write;
  int g1, g2;
  // counting the contacts that become constraint rows (constraint.py:
  // 155-156): dim 3 and inside the include margin; dim 0 counts none
  int dim = 0;
  float incl = 0.0f;
  int nact = 0;
  MJH_DEV void emit(float dist, V3 pos, const M33& frame, float margin) {
    if (!((dist - margin) < 0.0f)) return;
    if (dim == 3 && dist - incl < 0.0f) nact++;
    const int idx = base + k++;
    if (!write || idx >= cap) return;
    d->contact_dist[idx] = dist;
    st3(d->contact_pos + 3 * idx, pos);
    st9(d->contact_frame + 9 * idx, frame);
    d->contact_geom[2 * idx] = g1;
    d->contact_geom[2 * idx + 1] = g2;
    d->contact_worldid[idx] = world;
  }
};

MJH_DEV void plane_sphere_pt(V3 n, V3 ppos, V3 spos, float r, float* dist, V3* pos) {
  *dist = dot(spos - ppos, n) - r;
  *pos = spos - n * (r + 0.5f * (*dist));
}

MJH_DEV void sphere_sphere(V3 p1, float r1, V3 p2, float r2, float margin, ContactSink& s) {
  const V3 dir = p2 - p1;
  float dist = length(dir);
  const V3 n = (dist == 0.0f) ? V3{1.0f, 0.0f, 0.0f} : dir * (1.0f / dist);
  dist = dist - (r1 + r2);
  const V3 pos = p1 + n * (r1 + 0.5f * dist);
  s.emit(dist, pos, make_frame(n), margin);
}

// math.py:191-242 closest points between two segments
MJH_DEV V3 closest_segment_point(V3 a, V3 b, V3 pt) {
  const V3 ab = b - a;
  const float tt = dot(pt - a, ab) / (dot(ab, ab) + 1e-6f);
  return a + fminf(fmaxf(tt, 0.0f), 1.0f) * ab;
}
MJH_DEV void closest_segment_to_segment_points(V3 a0, V3 a1, V3 b0, V3 b1, V3* ra, V3* rb) {
  V3 da = a1 - a0, db = b1 - b0;
  float la = length(da), lb = length(db);
  if (la != 0.0f) da = da * (1.0f / la);
  if (lb != 0.0f) db = db * (1.0f / lb);
  const float ha = la * 0.5f, hb = lb * 0.5f;
  const V3 am = a0 + da * ha, bm = b0 + db * hb;
  const V3 tr = am - bm;
  const float dadb = dot(da, db), datr = dot(da, tr), dbtr = dot(db, tr);
  const float denom = 1.0f - dadb * dadb;
  const float ota = (-datr + dadb * dbtr) / (denom + 1e-6f);
  const float otb = dbtr + ota * dadb;
  const float ta = fminf(fmaxf(ota, -ha), ha);
  const float tb = fminf(fmaxf(otb, -hb), hb);
  const V3 ba = am + da * ta, bb = bm + db * tb;
  const V3 na = closest_segment_point(a0, a1, bb);
  const V3 nb = closest_segment_point(b0, b1, ba);
  const float d1 = dot(bb - na, bb - na), d2 = dot(ba - nb, ba - nb);
  if (d1 < d2) {
    *ra = na;
    *rb = bb;
  } else {
    *ra = ba;
    *rb = nb;
  }
}

// collision_primitive.py:235-268 dispatch on (type1 <= type2)
MJH_DEV void narrowphase_pair(const mjhModel& m, const float* gxpos, const float* gxmat, int g1, int g2,
                              ContactSink& s) {
  const int t1 = m.geom_type[g1], t2 = m.geom_type[g2];
  const float margin = fmaxf(m.geom_margin[g1], m.geom_margin[g2]);
  const V3 p1 = ld3(gxpos + 3 * g1), p2 = ld3(gxpos + 3 * g2);
  const M33 r1 = ld9(gxmat + 9 * g1), r2 = ld9(gxmat + 9 * g2);
  const float* s1 = m.geom_size + 3 * g1;
  const float* s2 = m.geom_size + 3 * g2;
  s.g1 = g1;
  s.g2 = g2;
  if (t1 == mjhGEOM_PLANE && t2 == mjhGEOM_SPHERE) {
    const V3 n = mcol(r1, 2);
    float dist;
    V3 pos;
    plane_sphere_pt(n, p1, p2, s2[0], &dist, &pos);
    s.emit(dist, pos, make_frame(n), margin);
  } else if (t1 == mjhGEOM_SPHERE && t2 == mjhGEOM_SPHERE) {
    sphere_sphere(p1, s1[0], p2, s2[0], margin, s);
  } else if (t1 == mjhGEOM_PLANE && t2 == mjhGEOM_CAPSULE) {
    const V3 n = mcol(r1, 2);
    const V3 axis = mcol(r2, 2);
    V3 b = axis - n * dot(n, axis);
    const float bn = length(b);
    if (bn != 0.0f) b = b * (1.0f / bn);
    if (bn < 0.5f) b = (-0.5f < n.y && n.y < 0.5f) ? V3{0.f, 1.f, 0.f} : V3{0.f, 0.f, 1.f};
    const V3 c = cross(n, b);
    M33 frame = {{n.x, n.y, n.z, b.x, b.y, b.z, c.x, c.y, c.z}};
    const V3 seg = axis * s2[1];
    float dist;
    V3 pos;
    plane_sphere_pt(n, p1, p2 + seg, s2[0], &dist, &pos);
    s.emit(dist, pos, frame, margin);
    plane_sphere_pt(n, p1, p2 - seg, s2[0], &dist, &pos);
    s.emit(dist, pos, frame, margin);
  } else if (t1 == mjhGEOM_PLANE && t2 == mjhGEOM_BOX) {
    const V3 n = mcol(r1, 2);
    const float dist = dot(p2 - p1, n);
    int count = 0;
    for (int i = 0; i < 8; i++) {
      V3 cl = V3{(i & 1) ? s2[0] : -s2[0], (i & 2) ? s2[1] : -s2[1], (i & 4) ? s2[2] : -s2[2]};
      const V3 corner = mulmv(r2, cl);
      const float ld = dot(n, corner);
      if (dist + ld > margin || ld > 0.0f) continue;
      const float cd = dist + ld;
      const V3 pos = corner + p2 + (n * cd) * (1.0f / -2.0f);
      s.emit(cd, pos, make_frame(n), margin);
      count++;
      if (count >= 4) break;
    }
  } else if (t1 == mjhGEOM_CAPSULE && t2 == mjhGEOM_CAPSULE) {
    const V3 a1 = mcol(r1, 2), a2 = mcol(r2, 2);
    const V3 seg1 = a1 * s1[1], seg2 = a2 * s2[1];
    V3 pt1, pt2;
    closest_segment_to_segment_points(p1 - seg1, p1 + seg1, p2 - seg2, p2 + seg2, &pt1, &pt2);
    sphere_sphere(pt1, s1[0], pt2, s2[0], margin, s);
  }
  // sphere-capsule and other pairs produce no contacts in the reference
}

// collision_driver.py:62-65: pairs are ordered by geom type
MJH_DEV void type_order(const mjhModel& m, int* g1, int* g2) {
  if (m.geom_type[*g1] > m.geom_type[*g2]) {
    const int tmp = *g1;
    *g1 = *g2;
    *g2 = tmp;
  }
}
// contact dimension and include margin of the pair's contacts, as
// contact_params mixes them (collision_driver.py:304-347)
MJH_DEV int pair_condim(const mjhModel& m, int g1, int g2) {
  const int p1 = m.geom_priority[g1], p2 = m.geom_priority[g2];
  const int c1 = m.geom_condim[g1], c2 = m.geom_condim[g2];
  return (p1 == p2) ? max(c1, c2) : (p1 > p2 ? c1 : c2);
}
MJH_DEV float pair_includemargin(const mjhModel& m, int g1, int g2) {
  return fmaxf(m.geom_margin[g1], m.geom_margin[g2]) - fmaxf(m.geom_gap[g1], m.geom_gap[g2]);
}

// collision_driver.py:304-347 contact parameter mixing of the contacts
// [idx, idx + n) of the geom pair (g1, g2)
MJH_DEV void contact_params(const mjhModel& m, const mjhData& d, int g1, int g2, int idx, int n) {
  const float incl = pair_includemargin(m, g1, g2);
  const float sm1 = m.geom_solmix[g1], sm2 = m.geom_solmix[g2];
  float mix = sm1 / (sm1 + sm2);
  if (sm1 < MJH_MINVAL && sm2 < MJH_MINVAL) mix = 0.5f;
  else if (sm1 < MJH_MINVAL && sm2 >= MJH_MINVAL) mix = 0.0f;
  else if (sm1 >= MJH_MINVAL && sm2 < MJH_MINVAL) mix = 1.0f;
  const int p1 = m.geom_priority[g1], p2 = m.geom_priority[g2];
  if (p1 != p2) mix = p1 > p2 ? 1.0f : 0.0f;
  const int dim = pair_condim(m, g1, g2);
  const float* sr1 = m.geom_solref + 2 * g1;
  const float* sr2 = m.geom_solref + 2 * g2;
  float sr[2];
  if (sr1[0] > 0.0f && sr2[0] > 0.0f) {
    sr[0] = mix * sr1[0] + (1.0f - mix) * sr2[0];
    sr[1] = mix * sr1[1] + (1.0f - mix) * sr2[1];
  } else {
    sr[0] = fminf(sr1[0], sr2[0]);
    sr[1] = fminf(sr1[1], sr2[1]);
  }
  const float* f1 = m.geom_friction + 3 * g1;
  const float* f2 = m.geom_friction + 3 * g2;
  const float fa = fmaxf(f1[0], f2[0]), fb = fmaxf(f1[1], f2[1]), fc = fmaxf(f1[2], f2[2]);
  const float* si1 = m.geom_solimp + 5 * g1;
  const float* si2 = m.geom_solimp + 5 * g2;
  float si[5];
#pragma unroll
  for (int k = 0; k < 5; k++) si[k] = mix * si1[k] + (1.0f - mix) * si2[k];
  for (int c = idx; c < idx + n; c++) {
    d.contact_dim[c] = dim;
    d.contact_solref[2 * c] = sr[0];
    d.contact_solref[2 * c + 1] = sr[1];
    d.contact_includemargin[c] = incl;
    float* fr = d.contact_friction + 5 * c;
    fr[0] = fa;
    fr[1] = fa;
    fr[2] = fb;
    fr[3] = fc;
    fr[4] = fc;
#pragma unroll
    for (int k = 0; k < 5; k++) d.contact_solimp[5 * c + k] = si[k];
  }
}

// collision_primitive.py:271-274 primitive_narrowphase over this world's
// broadphase block, then the contact parameters of the contacts written
template <int TEAM>
MJH_DEV void narrowphase(const mjhModel& m, const mjhData& d, const Team<TEAM>& t) {
  const int w = t.world, ng = m.ngeom;
  int padr = 0, npair = 0;
  if (t.valid) {
    padr = d.world_pair[2 * w];
    npair = d.world_pair[2 * w + 1];
  }
  const float* gxpos = WPTR(d.geom_xpos, w, ng * 3);
  const float* gxmat = WPTR(d.geom_xmat, w, ng * 9);
  // pass 1 counts the contacts of each pair; the first pair of each lane
  // is kept in registers for the write pass
  int count = 0;
  int c0 = 0, g01 = 0, g02 = 0;
  for (int p = t.lane; p < npair; p += TEAM) {
    const int g1 = d.collision_pair[2 * (padr + p)], g2 = d.collision_pair[2 * (padr + p) + 1];
    ContactSink s{&d, 0, 0, 0, w, false, 0, 0};
    narrowphase_pair(m, gxpos, gxmat, g1, g2, s);
    count += s.k;
    if (p == t.lane) {
      c0 = s.k;
      g01 = g1;
      g02 = g2;
    }
  }
  const int total = team_sum_int<TEAM>(count);
  const int base = team_reserve<TEAM>(d.ncon, t.valid ? total : 0);
  const int cap = d.nconmax;
  const int nwritten = total <= 0 ? 0 : max(0, min(total, cap - base));
  if (t.valid && t.lane == 0) {
    d.world_con[2 * w] = base;
    d.world_con[2 * w + 1] = nwritten;
    if (base + total > cap) atomicAdd(d.overflow + 1, min(total, base + total - cap));
  }
  int k = 0;
  for (int p0 = 0; p0 < npair; p0 += TEAM) {
    const int p = p0 + t.lane;
    int c = 0;
    int g1 = 0, g2 = 0;
    if (p0 == 0) {
      c = c0;
      g1 = g01;
      g2 = g02;
    } else if (p < npair) {
      g1 = d.collision_pair[2 * (padr + p)];
      g2 = d.collision_pair[2 * (padr + p) + 1];
      ContactSink s{&d, 0, 0, 0, w, false, 0, 0};
      narrowphase_pair(m, gxpos, gxmat, g1, g2, s);
      c = s.k;
    }
    // a pair yields at most 4 contacts (plane-box)
    int tot;
    const int off = team_exscan_bits<TEAM, 3>(c, &tot);
    if (c > 0) {
      ContactSink s{&d, base, cap, k + off, w, true, 0, 0};
      narrowphase_pair(m, gxpos, gxmat, g1, g2, s);
      // contacts of this pair that fit in the pool
      const int n = min(base + k + off + c, cap) - (base + k + off);
      if (n > 0) contact_params(m, d, g1, g2, base + k + off, n);
    }
    k += tot;
  }
  sync();
}

// collision_driver.py:525-547 collision (counters are reset by the caller)
template <int TEAM>
MJH_DEV void collision(const mjhModel& m, const mjhData& d, float* lds, const Team<TEAM>& t) {
  if (m.ngeom <= 100) nxn_broadphase(m, d, lds, t);
  else sap_broadphase(m, d, lds, t);
  sync();
  narrowphase(m, d, t);
}

}  // namespace mjh
