// Icosahedral grid generation for FIM, host C++.
//
// Restates the reference prep chain for an all-bisection grid with the
// modified recursive algorithm (gtype=2) in IJ order (curve=0):
//   grid   FIMsrc/prep/grid/icos.F90:27-239 with top_gridpoints.F90,
//          rotate.F90, top_triangles.F90, bisect_triangle.F90:254-843,
//          middle.F90, third.F90, avg.F90
//   ginfo  FIMsrc/prep/grid/GridGen.F90:1-497 (pattern neighbours, Voronoi
//          corners) and perm.F90:44-54 (identity permutation)
//   getlvl FIMsrc/prep/grid/getlvl.F90:2-371 (proxs and the FP32 metrics)
// Mixed-precision details of the reference are kept on purpose, because they
// decide the bits of the result: pi computed as 4*atan(1) in single precision
// and widened, in-place degree<->radian round trips of the callers'
// arguments, single-precision 1/3, and FP32 metrics with the C library's
// single-precision sin/cos.  Build without fp contraction or fast math.
//
// hfcoef (getlvl.F90:219-269) is the least-squares edge reconstruction
// weight; its pseudo-inverse is the single-precision SVD of lsq_svd.cpp,
// which follows the LAPACK routines the reference vendors, so hfcoef is the
// reference's bit for bit as well.
#include <math.h>
#include <string.h>

#include <vector>

#include "fimhip.h"

namespace {

struct GP {
  double ll[2];   // lat, lon (degrees while building)
  int sn;
};

struct GPWnb {
  double ll[2];
  double nb[6][2];   // Neighbor(1:2, j) -> nb[j][0] lat, nb[j][1] lon
  int sn;
};

struct Tri {
  GP v[3];
  int start_idx, end_idx;
};

// 4.0*ATAN(1.0) in default (single) precision, widened to REAL*8
static double pi_single() {
  float p = 4.0f * atanf(1.0f);
  return (double)p;
}

static inline double dot3(const double *a, const double *b) {
  double s = 0.;
  s = s + a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}

// middle.F90:23-71 (arguments modified in place, as in the reference)
static void middle(double *p1, double *p2, double *p) {
  const double pi = pi_single();
  const double d2r = pi / 180.0;
  p1[0] = p1[0] * d2r; p1[1] = p1[1] * d2r;
  p2[0] = p2[0] * d2r; p2[1] = p2[1] * d2r;
  double x1[3] = {cos(p1[0]) * cos(p1[1]), cos(p1[0]) * sin(p1[1]), sin(p1[0])};
  double x2[3] = {cos(p2[0]) * cos(p2[1]), cos(p2[0]) * sin(p2[1]), sin(p2[0])};
  double x[3];
  for (int i = 0; i < 3; i++) x[i] = 0.5 * (x1[i] + x2[i]);
  double nrm = sqrt(dot3(x, x));
  for (int i = 0; i < 3; i++) x[i] = x[i] / nrm;
  p[0] = atan2(x[2], sqrt(x[0] * x[0] + x[1] * x[1]));
  p[1] = atan2(x[1], x[0]);
  if (p[1] < -pi / 2.0) p[1] = p[1] + 2 * pi;
  p1[0] = p1[0] / d2r; p1[1] = p1[1] / d2r;
  p2[0] = p2[0] / d2r; p2[1] = p2[1] / d2r;
  p[0] = p[0] / d2r; p[1] = p[1] / d2r;
}

// third.F90:23-72
static void one_third(double *p1, double *p2, double *p) {
  const double pi = pi_single();
  const double d2r = pi / 180.0;
  p1[0] = p1[0] * d2r; p1[1] = p1[1] * d2r;
  p2[0] = p2[0] * d2r; p2[1] = p2[1] * d2r;
  double x1[3] = {cos(p1[0]) * cos(p1[1]), cos(p1[0]) * sin(p1[1]), sin(p1[0])};
  double x2[3] = {cos(p2[0]) * cos(p2[1]), cos(p2[0]) * sin(p2[1]), sin(p2[0])};
  const double f = (double)(1.0f / 3.0f);
  double theta = acos(dot3(x1, x2));
  double c1 = sin((1 - f) * theta) / sin(theta);
  double c2 = sin(f * theta) / sin(theta);
  double x[3];
  for (int i = 0; i < 3; i++) x[i] = c1 * x1[i] + c2 * x2[i];
  p[0] = atan2(x[2], sqrt(x[0] * x[0] + x[1] * x[1]));
  p[1] = atan2(x[1], x[0]);
  if (p[1] < -pi / 2.0) p[1] = p[1] + 2 * pi;
  p1[0] = p1[0] / d2r; p1[1] = p1[1] / d2r;
  p2[0] = p2[0] / d2r; p2[1] = p2[1] / d2r;
  p[0] = p[0] / d2r; p[1] = p[1] / d2r;
}

// avg.F90:1-53
static void average(double *p1, double *p2, double *p3, double *p) {
  const double pi = pi_single();
  const double d2r = pi / 180.0;
  p1[0] = p1[0] * d2r; p1[1] = p1[1] * d2r;
  p2[0] = p2[0] * d2r; p2[1] = p2[1] * d2r;
  p3[0] = p3[0] * d2r; p3[1] = p3[1] * d2r;
  double x1[3] = {cos(p1[0]) * cos(p1[1]), cos(p1[0]) * sin(p1[1]), sin(p1[0])};
  double x2[3] = {cos(p2[0]) * cos(p2[1]), cos(p2[0]) * sin(p2[1]), sin(p2[0])};
  double x3[3] = {cos(p3[0]) * cos(p3[1]), cos(p3[0]) * sin(p3[1]), sin(p3[0])};
  double x[3];
  for (int i = 0; i < 3; i++) x[i] = (x1[i] + x2[i] + x3[i]) / 3.0;
  double nrm = sqrt(dot3(x, x));
  for (int i = 0; i < 3; i++) x[i] = x[i] / nrm;
  p[0] = atan2(x[2], sqrt(x[0] * x[0] + x[1] * x[1]));
  p[1] = atan2(x[1], x[0]);
  if (p[1] < -pi / 2.0) p[1] = p[1] + 2 * pi;
  p1[0] = p1[0] / d2r; p1[1] = p1[1] / d2r;
  p2[0] = p2[0] / d2r; p2[1] = p2[1] / d2r;
  p3[0] = p3[0] / d2r; p3[1] = p3[1] / d2r;
  p[0] = p[0] / d2r; p[1] = p[1] / d2r;
}

// rotate.F90:8-92
static void rotate_latlon(double *ll, double sin_t, double cos_t, double sin_l, double cos_l) {
  const double pi = pi_single();
  double x[3] = {cos(ll[0]) * cos(ll[1]), cos(ll[0]) * sin(ll[1]), sin(ll[0])};
  double x1[3] = {cos_t * x[0] - sin_t * x[1], sin_t * x[0] + cos_t * x[1], x[2]};
  double c11 = cos_l + (1 - cos_l) * sin_t * sin_t;
  double c12 = -((1 - cos_l) * sin_t * cos_t);
  double c13 = -(sin_l * cos_t);
  double c21 = c12;
  double c22 = cos_l + (1 - cos_l) * cos_t * cos_t;
  double c23 = -(sin_l * sin_t);
  double c31 = -c13, c32 = -c23, c33 = cos_l;
  double x2[3] = {c11 * x1[0] + c12 * x1[1] + c13 * x1[2], c21 * x1[0] + c22 * x1[1] + c23 * x1[2],
                  c31 * x1[0] + c32 * x1[1] + c33 * x1[2]};
  ll[0] = atan2(x2[2], sqrt(x2[0] * x2[0] + x2[1] * x2[1]));
  ll[1] = atan2(x2[1], x2[0]);
  if (ll[1] < -pi / 2.0) ll[1] = ll[1] + 2 * pi;
}

static void rotate(double theta, double lambda, GPWnb *top, int n) {
  const double pi = pi_single();
  const double d2r = pi / 180.0;
  theta = theta * d2r;
  lambda = lambda * d2r;
  double sin_t = sin(theta), cos_t = cos(theta), sin_l = sin(lambda), cos_l = cos(lambda);
  for (int i = 0; i < n; i++) {
    double ll[2] = {top[i].ll[0] * d2r, top[i].ll[1] * d2r};
    rotate_latlon(ll, sin_t, cos_t, sin_l, cos_l);
    top[i].ll[0] = ll[0] / d2r;
    top[i].ll[1] = ll[1] / d2r;
    for (int j = 0; j < 6; j++) {
      double q[2] = {top[i].nb[j][0] * d2r, top[i].nb[j][1] * d2r};
      rotate_latlon(q, sin_t, cos_t, sin_l, cos_l);
      top[i].nb[j][0] = q[0] / d2r;
      top[i].nb[j][1] = q[1] / d2r;
    }
  }
}

// top_gridpoints.F90:8-108 (top[] is 1-based through index i-1)
static void set_top_gridpoints(GPWnb *top) {
  const double lat0 = (double)26.5650428671016f, lon0 = -72.0;
  for (int i = 0; i < 12; i++) { top[i].nb[5][0] = 0.; top[i].nb[5][1] = 0.; }
  top[0].ll[0] = 90.0; top[0].ll[1] = 0.0;
  for (int i = 1; i <= 5; i++) {
    top[0].nb[i - 1][0] = lat0;
    top[0].nb[i - 1][1] = lon0 + (double)(72.0f * (float)i);
  }
  for (int i = 2; i <= 6; i++) {
    GPWnb &t = top[i - 1];
    t.ll[0] = top[0].nb[i - 2][0];
    t.ll[1] = top[0].nb[i - 2][1];
    t.nb[0][0] = top[0].ll[0]; t.nb[0][1] = top[0].ll[1];
    int n = i - 1; if (n == 1) n = 5;
    t.nb[1][0] = top[0].nb[n - 1][0]; t.nb[1][1] = top[0].nb[n - 1][1];
    n = i + 1; if (n == 7) n = 2;
    t.nb[2][0] = top[0].nb[n - 1][0]; t.nb[2][1] = top[0].nb[n - 1][1];
  }
  top[6].ll[0] = -90.0; top[6].ll[1] = 0.0;
  for (int i = 1; i <= 5; i++) {
    top[6].nb[i - 1][0] = -lat0;
    top[6].nb[i - 1][1] = lon0 + 36.0 + (double)(72.0f * (float)i);
  }
  for (int i = 8; i <= 12; i++) {
    GPWnb &t = top[i - 1];
    t.ll[0] = top[6].nb[i - 8][0];
    t.ll[1] = top[6].nb[i - 8][1];
    t.nb[0][0] = top[6].ll[0]; t.nb[0][1] = top[6].ll[1];
    int n = i - 1; if (n == 7) n = 12;
    t.nb[1][0] = top[6].nb[n - 8][0]; t.nb[1][1] = top[6].nb[n - 8][1];
    n = i; if (n == 12) n = 8;
    t.nb[2][0] = top[6].nb[n - 8][0]; t.nb[2][1] = top[6].nb[n - 8][1];
  }
  for (int i = 2; i <= 6; i++) {
    int n = i + 6;
    top[i - 1].nb[3][0] = top[n - 1].ll[0]; top[i - 1].nb[3][1] = top[n - 1].ll[1];
    top[n - 1].nb[3][0] = top[i - 1].ll[0]; top[n - 1].nb[3][1] = top[i - 1].ll[1];
    n = i + 5; if (n == 7) n = 12;
    top[i - 1].nb[4][0] = top[n - 1].ll[0]; top[i - 1].nb[4][1] = top[n - 1].ll[1];
    top[n - 1].nb[4][0] = top[i - 1].ll[0]; top[n - 1].nb[4][1] = top[i - 1].ll[1];
  }
}

// top_triangles.F90:8-84
static void set_top_triangle(Tri *high, const GPWnb *top) {
  for (int i = 1; i <= 5; i++) {
    Tri &h = high[i - 1];
    h.v[0].ll[0] = top[0].ll[0]; h.v[0].ll[1] = top[0].ll[1];
    h.v[1].ll[0] = top[0].nb[i - 1][0]; h.v[1].ll[1] = top[0].nb[i - 1][1];
    int n = i + 1; if (n == 6) n = 1;
    h.v[2].ll[0] = top[0].nb[n - 1][0]; h.v[2].ll[1] = top[0].nb[n - 1][1];
  }
  for (int i = 6; i <= 10; i++) {
    Tri &h = high[i - 1];
    h.v[0].ll[0] = top[6].ll[0]; h.v[0].ll[1] = top[6].ll[1];
    h.v[2].ll[0] = top[6].nb[i - 6][0]; h.v[2].ll[1] = top[6].nb[i - 6][1];
    int n = i + 1; if (n == 11) n = 6;
    h.v[1].ll[0] = top[6].nb[n - 6][0]; h.v[1].ll[1] = top[6].nb[n - 6][1];
  }
  for (int i = 1; i <= 5; i++) {
    Tri &h = high[i + 9];
    h.v[0].ll[0] = top[i].ll[0]; h.v[0].ll[1] = top[i].ll[1];
    h.v[2].ll[0] = top[i].nb[3][0]; h.v[2].ll[1] = top[i].nb[3][1];
    h.v[1].ll[0] = top[i].nb[4][0]; h.v[1].ll[1] = top[i].nb[4][1];
  }
  for (int i = 8; i <= 12; i++) {
    Tri &h = high[i + 7];
    h.v[0].ll[0] = top[i - 1].ll[0]; h.v[0].ll[1] = top[i - 1].ll[1];
    h.v[2].ll[0] = top[i - 1].nb[3][0]; h.v[2].ll[1] = top[i - 1].nb[3][1];
    h.v[1].ll[0] = top[i - 1].nb[4][0]; h.v[1].ll[1] = top[i - 1].nb[4][1];
  }
}

struct Builder {
  int ml = 0, sl = 1, offset = 0;
  int init_a_sn = 0, init_b_sn = 0, init_c_sn = 0;
  std::vector<GP> gp;   // 1-based

  // bisect_triangle.F90:709-721 (REAL function)
  float row(int x) const {
    int s = init_c_sn - x;
    int n = (int)((sqrtf((float)(1 + 8 * s)) - 1.0f) / 2.0f);
    return (float)(sl - n);
  }

  // bisect_triangle.F90:685-707
  void compute_index(int a_sn, int b_sn, int c_sn, int *a_b_sn, int *a_c_sn, int *b_c_sn, int delta) const {
    int init_stride = (int)((float)(sl + 1) - row(a_sn));
    int nrtm = (int)((row(c_sn) - row(a_sn)) / 2.0f);
    int stride1 = (2 * init_stride - (nrtm - 1)) * nrtm / 2;
    int stride2 = (2 * (init_stride - 1) - (nrtm - 1)) * nrtm / 2;
    *a_c_sn = a_sn + stride1;
    if (delta == 1) {
      *a_b_sn = (a_sn + b_sn) / 2;
      *b_c_sn = b_sn + stride2;
    } else {
      *b_c_sn = (b_sn + c_sn) / 2;
      *a_b_sn = a_sn + stride2;
    }
  }

  void store(const GP &p) {
    GP &q = gp[p.sn + offset];
    q.ll[0] = p.ll[0];
    q.ll[1] = p.ll[1];
    q.sn = p.sn;
  }

  // bisect_triangle.F90:254-338 (delta = 1) and 340-424 (reversed, delta = 0)
  void bisect(GP &a, GP &b, GP &c, GP &a1, GP &a2, GP &a3, int level, int side, bool rev) {
    int clevel = level + 1;
    store(a);
    store(b);
    store(c);
    if (clevel > ml) return;
    GP a_b{}, a_c{}, b_c{}, a_a_b{}, a_b_b{}, a_a_c{}, a_c_c{}, b_b_c{}, b_c_c{};
    GP anc1{}, anc2{}, anc3{}, t1{}, t2{}, t3{};
    compute_index(a.sn, b.sn, c.sn, &a_b.sn, &a_c.sn, &b_c.sn, rev ? 0 : 1);
    middle(a.ll, b.ll, a_b.ll);
    middle(a.ll, c.ll, a_c.ll);
    middle(b.ll, c.ll, b_c.ll);
    if (side == 1) {
      one_third(b_c.ll, a1.ll, t1.ll);
      one_third(a_c.ll, a2.ll, t2.ll);
      middle(a.ll, b.ll, t3.ll);
      average(t1.ll, t2.ll, t3.ll, a_b.ll);
      a3 = a_b;
    } else if (side == 2) {
      one_third(b_c.ll, a1.ll, t1.ll);
      one_third(a_b.ll, a2.ll, t2.ll);
      middle(a.ll, c.ll, t3.ll);
      average(t1.ll, t2.ll, t3.ll, a_c.ll);
      a3 = a_c;
    } else if (side == 3) {
      one_third(a_c.ll, a1.ll, t1.ll);
      one_third(a_b.ll, a2.ll, t2.ll);
      middle(b.ll, c.ll, t3.ll);
      average(t1.ll, t2.ll, t3.ll, b_c.ll);
      a3 = b_c;
    } else if (side == 4) {
      a_b = a1;
      a_c = a2;
      b_c = a3;
    }
    middle(a.ll, a_b.ll, a_a_b.ll);
    middle(a_b.ll, b.ll, a_b_b.ll);
    middle(a.ll, a_c.ll, a_a_c.ll);
    middle(a_c.ll, c.ll, a_c_c.ll);
    middle(b.ll, b_c.ll, b_b_c.ll);
    middle(b_c.ll, c.ll, b_c_c.ll);
    // all refinements are bisections (SubdivNum = 2)
    bisect(a, a_b, a_c, b_b_c, b_c_c, anc1, clevel, 3, rev);
    bisect(a_b, b, b_c, a_a_c, a_c_c, anc2, clevel, 2, rev);
    bisect(a_c, b_c, c, a_a_b, a_b_b, anc3, clevel, 1, rev);
    bisect(a_b, a_c, b_c, anc1, anc2, anc3, clevel, 4, !rev);
  }

  // bisect_triangle.F90:723-794
  void remove_edge1(int start_idx, int &end_idx) {
    int stride = sl + 1, idx = start_idx;
    while (idx <= end_idx && stride >= 0) {
      gp[idx].sn = -1;
      idx = idx + stride;
      stride = stride - 1;
    }
    stride = 0;
    int idx2 = start_idx;
    for (idx = start_idx; idx <= end_idx; idx++) {
      if (gp[idx].sn == -1) { stride++; continue; }
      gp[idx2].ll[0] = gp[idx].ll[0];
      gp[idx2].ll[1] = gp[idx].ll[1];
      gp[idx2].sn = gp[idx].sn - stride;
      idx2++;
    }
    end_idx = end_idx - stride;
  }

  void remove_edge2(int start_idx, int &end_idx) {
    for (int idx = start_idx; idx <= start_idx + sl; idx++) gp[idx].sn = -1;
    int stride = sl - 1, idx = start_idx + sl * 2;
    while (idx <= end_idx && stride >= 0) {
      gp[idx].sn = -1;
      idx = idx + stride;
      stride = stride - 1;
    }
    stride = 0;
    int idx2 = start_idx;
    for (idx = start_idx; idx <= end_idx; idx++) {
      if (gp[idx].sn == -1) { stride++; continue; }
      gp[idx2].ll[0] = gp[idx].ll[0];
      gp[idx2].ll[1] = gp[idx].ll[1];
      gp[idx2].sn = gp[idx].sn - stride;
      idx2++;
    }
    end_idx = end_idx - stride;
  }

  // bisect_triangle.F90:796-842
  void combine(Tri &t1, Tri &t2) {
    const int last = gp[t2.end_idx].sn;
    for (int idx = t2.start_idx; idx <= t2.end_idx; idx++) gp[idx].sn = last - gp[idx].sn;
    // the reference reads gp(end_idx)%seqnum after earlier entries changed; the
    // end entry itself is updated last, so 'last' is the original value
    int stride = sl - 3, gap = 1, delta = 1;
    int idx = t1.start_idx + (2 * sl - 1);
    while (idx <= t1.end_idx) {
      int idx2;
      for (idx2 = idx; idx2 <= idx + stride; idx2++) gp[idx2].sn = gp[idx2].sn + gap;
      stride--;
      delta++;
      gap += delta;
      idx = idx2;
    }
    stride = 0;
    gap = 2 * sl - 1;
    delta = sl - 1;
    idx = t2.end_idx;
    while (idx >= t2.start_idx) {
      int idx2;
      for (idx2 = idx; idx2 >= idx - stride; idx2--) gp[idx2].sn = gp[idx2].sn + gap;
      stride++;
      delta--;
      gap += delta;
      idx = idx2;
    }
  }
};

// ---------------------------------------------------------------------------
// GridGen.F90:504-549 and 552-640
static int grid_point_type(int seq, int diam_sz, int edge_ln) {
  int i = (seq - 1) % (2 * diam_sz);
  if (i < edge_ln) return 1;
  if (i % edge_ln == 0 && i <= diam_sz) return 2;
  if (i % edge_ln == 0 && i < 2 * diam_sz) return 3;
  if (i % edge_ln == 1 && i < diam_sz) return 4;
  if (i % edge_ln == 1 && i < 2 * diam_sz) return 5;
  if (i >= 2 * diam_sz - edge_ln + 1) return 6;
  return 0;
}

static int special_points(int i, int *nb, int diam_sz, int edge_ln) {
  int type = 0;
  int r_i = (i - 1) % (2 * diam_sz);
  int pg_idx;
  if (i == 1) {
    nb[0] = 2;
    for (int k = 2; k <= 5; k++) nb[k - 1] = nb[k - 2] + 2 * diam_sz;
    nb[5] = -1;
    type = 1;
  } else if (i == 10 * diam_sz + 2) {
    nb[0] = 10 * diam_sz + 1;
    for (int k = 2; k <= 5; k++) nb[k - 1] = nb[k - 2] - 2 * diam_sz;
    nb[5] = -1;
    type = 1;
  } else if (r_i == 1) {
    nb[0] = 1;
    nb[1] = (i - 2 * diam_sz) > 0 ? i - 2 * diam_sz : i + 8 * diam_sz;
    nb[2] = nb[1] + edge_ln;
    nb[3] = i + 1;
    nb[4] = i + edge_ln;
    nb[5] = (i + 2 * diam_sz) % (10 * diam_sz);
    type = 2;
  } else if (r_i == 0) {
    nb[0] = i - edge_ln;
    nb[1] = (i - 4 * diam_sz) > 0 ? i - 2 * diam_sz - 1 : i + 8 * diam_sz - 1;
    nb[2] = nb[1] + 1;
    nb[3] = 10 * diam_sz + 2;
    nb[4] = (i + 2 * diam_sz) <= 10 * diam_sz + 1 ? i + 2 * diam_sz : i + 2 * diam_sz - 10 * diam_sz;
    nb[5] = i - 1;
    type = 2;
  } else if (r_i == edge_ln) {
    nb[0] = i - 1;
    pg_idx = i / (2 * diam_sz);
    pg_idx = pg_idx >= 1 ? pg_idx - 1 : 4;
    nb[1] = pg_idx * 2 * diam_sz + diam_sz - edge_ln + 2;
    nb[2] = nb[1] + edge_ln;
    nb[3] = i + edge_ln;
    nb[4] = nb[3] - 1;
    nb[5] = -1;
    type = 1;
  } else if (r_i == diam_sz + edge_ln) {
    nb[0] = i - 1;
    nb[1] = i - edge_ln;
    pg_idx = i / (2 * diam_sz);
    pg_idx = pg_idx >= 1 ? pg_idx - 1 : 4;
    nb[2] = pg_idx * 2 * diam_sz + 2 * diam_sz - edge_ln + 2;
    nb[3] = i + edge_ln;
    nb[4] = nb[3] - 1;
    nb[5] = -1;
    type = 1;
  } else if (r_i == 2 * diam_sz - edge_ln + 1) {
    nb[0] = i - edge_ln;
    nb[1] = nb[0] + 1;
    nb[2] = nb[1] + edge_ln;
    pg_idx = (i - 1) / (2 * diam_sz);
    pg_idx = pg_idx < 4 ? pg_idx + 1 : 0;
    r_i = (i - 1) % (2 * diam_sz);
    nb[3] = pg_idx * 2 * diam_sz + r_i - diam_sz + 3 * edge_ln;
    nb[4] = nb[3] - edge_ln;
    nb[5] = nb[4] - edge_ln;
    type = 2;
  }
  return type;
}

static void pattern_neighbours(int i, int *nb, int diam_sz, int edge_ln) {
  int gp_type = grid_point_type(i, diam_sz, edge_ln);
  int pg_idx, r_i;
  switch (gp_type) {
    case 0:
      nb[0] = i - 1; nb[1] = i - edge_ln; nb[2] = i - edge_ln + 1;
      nb[3] = i + 1; nb[4] = i + edge_ln; nb[5] = i + edge_ln - 1;
      break;
    case 1:
      nb[0] = i - 1;
      pg_idx = (i - 1) / (2 * diam_sz);
      pg_idx = pg_idx >= 1 ? pg_idx - 1 : 4;
      r_i = (i - 1) % (2 * diam_sz);
      nb[1] = pg_idx * 2 * diam_sz + (r_i - 1) * edge_ln + 2;
      nb[2] = nb[1] + edge_ln;
      nb[3] = i + 1; nb[4] = i + edge_ln; nb[5] = i + edge_ln - 1;
      break;
    case 2:
      nb[0] = i - edge_ln;
      pg_idx = (i - 1) / (2 * diam_sz);
      pg_idx = pg_idx >= 1 ? pg_idx - 1 : 4;
      r_i = (i - 1) % (2 * diam_sz);
      nb[1] = pg_idx * 2 * diam_sz + diam_sz - 2 * edge_ln + r_i + 2;
      nb[2] = nb[1] + edge_ln;
      nb[3] = i + edge_ln; nb[4] = nb[3] - 1; nb[5] = i - 1;
      break;
    case 3:
      nb[0] = i - edge_ln;
      pg_idx = (i - 1) / (2 * diam_sz);
      pg_idx = pg_idx >= 1 ? pg_idx - 1 : 4;
      r_i = (i - 1) % (2 * diam_sz);
      nb[1] = (pg_idx + 1) * 2 * diam_sz - edge_ln + (r_i - (diam_sz + 2 * edge_ln)) / edge_ln + 2;
      nb[2] = nb[1] + 1;
      nb[3] = i + edge_ln; nb[4] = nb[3] - 1; nb[5] = i - 1;
      break;
    case 4:
      nb[0] = i - edge_ln; nb[1] = nb[0] + 1; nb[2] = i + 1; nb[3] = i + edge_ln;
      pg_idx = (i - 1) / (2 * diam_sz);
      pg_idx = pg_idx < 4 ? pg_idx + 1 : 0;
      r_i = (i - 1) % (2 * diam_sz);
      nb[4] = pg_idx * 2 * diam_sz + 2 + ((r_i - 1) / edge_ln);
      nb[5] = nb[4] - 1;
      break;
    case 5:
      nb[0] = i - edge_ln; nb[1] = nb[0] + 1; nb[2] = nb[1] + edge_ln; nb[3] = i + edge_ln;
      pg_idx = (i - 1) / (2 * diam_sz);
      pg_idx = pg_idx < 4 ? pg_idx + 1 : 0;
      r_i = (i - 1) % (2 * diam_sz);
      nb[4] = pg_idx * 2 * diam_sz + r_i - diam_sz + 2 * edge_ln;
      nb[5] = nb[4] - edge_ln;
      break;
    case 6:
      nb[0] = i - 1; nb[1] = i - edge_ln; nb[2] = nb[1] + 1; nb[3] = i + 1;
      pg_idx = (i - 1) / (2 * diam_sz);
      pg_idx = pg_idx < 4 ? pg_idx + 1 : 0;
      r_i = (i - 1) % (2 * diam_sz);
      nb[4] = pg_idx * 2 * diam_sz + diam_sz + (r_i - 2 * diam_sz + edge_ln + 1) * edge_ln + 1;
      nb[5] = nb[4] - edge_ln;
      break;
  }
}

// getlvl.F90:385-396
static void ll2xy(float lonc, float latc, float lon, float lat, float *xm, float *ym) {
  float mf = 2.0f / (1.0f + sinf(lat) * sinf(latc) + cosf(lat) * cosf(latc) * cosf(lon - lonc));
  *xm = mf * (cosf(lat) * sinf(lon - lonc));
  *ym = mf * ((sinf(lat) * cosf(latc) - cosf(lat) * sinf(latc) * cosf(lon - lonc)));
}

// getlvl.F90:398-425
static void middle_r(const float *p1, const float *p2, float *p) {
  const float pi = atanf(1.0f) * 4.0f;
  float x1[3] = {cosf(p1[0]) * cosf(p1[1]), cosf(p1[0]) * sinf(p1[1]), sinf(p1[0])};
  float x2[3] = {cosf(p2[0]) * cosf(p2[1]), cosf(p2[0]) * sinf(p2[1]), sinf(p2[0])};
  float x[3];
  for (int i = 0; i < 3; i++) x[i] = 0.5f * (x1[i] + x2[i]);
  float nrm = sqrtf(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  for (int i = 0; i < 3; i++) x[i] = x[i] / nrm;
  p[0] = atan2f(x[2], sqrtf(x[0] * x[0] + x[1] * x[1]));
  p[1] = atan2f(x[1], x[0]);
  if (p[1] < 0.0f) p[1] = p[1] + 2 * pi;
}

}  // namespace

// ---------------------------------------------------------------------------
static int icos_grid_impl(int glvl, float *lat_out, float *lon_out, int *nprox_out, int *prox_out,
                          int *proxs_out, float *area_out, float *cs_out, float *sn_out, float *svc_out,
                          float *sve_out, float *sideln_out, float *rprox_out, float *corner_out,
                          float *hfcoef_out, int *inv_perm_out, float *edge_out) {
  if (glvl < 0 || glvl > 10) return FIMHIP_EINVAL;
  // ---- icos.F90: grid points -------------------------------------------
  GPWnb top[12];
  memset(top, 0, sizeof top);
  set_top_gridpoints(top);
  rotate(10.0, 0.0, top, 12);
  Tri tri[20];
  memset(tri, 0, sizeof tri);
  set_top_triangle(tri, top);
  Builder B;
  B.sl = 1;
  for (int i = 0; i < glvl; i++) B.sl *= 2;
  B.ml = glvl;
  const int sl = B.sl;
  const int buf_sz = 20 * (sl + 2) * (sl + 1) / 2;
  B.gp.assign(buf_sz + 2, GP{{0., 0.}, 0});
  const int nip = 10 * sl * sl + 2;
  B.init_a_sn = 0;
  B.init_b_sn = sl;
  B.init_c_sn = (sl + 2) * (sl + 1) / 2 - 1;
  for (int i = 0; i < 20; i++) {       // comp_grid, icos.F90:127-168
    B.offset = i * (B.init_c_sn + 1) + 1;
    tri[i].v[0].sn = B.init_a_sn;
    tri[i].v[1].sn = B.init_b_sn;
    tri[i].v[2].sn = B.init_c_sn;
    tri[i].start_idx = B.offset + B.init_a_sn;
    tri[i].end_idx = B.offset + B.init_c_sn;
    GP ph{};
    B.bisect(tri[i].v[0], tri[i].v[1], tri[i].v[2], ph, ph, ph, 0, 0, false);
  }
  // proc_grid, icos.F90:171-238
  for (int i = 1; i <= 5; i++) B.remove_edge1(tri[i - 1].start_idx, tri[i - 1].end_idx);
  for (int i = 11; i <= 15; i++) B.remove_edge1(tri[i - 1].start_idx, tri[i - 1].end_idx);
  for (int i = 6; i <= 10; i++) B.remove_edge2(tri[i - 1].start_idx, tri[i - 1].end_idx);
  for (int i = 16; i <= 20; i++) B.remove_edge2(tri[i - 1].start_idx, tri[i - 1].end_idx);
  for (int i = 1; i <= 5; i++) B.combine(tri[i - 1], tri[i + 14]);
  for (int i = 6; i <= 9; i++) B.combine(tri[i + 5], tri[i - 1]);
  B.combine(tri[10], tri[9]);
  const int off = sl * sl;
  auto shift = [&](int j, int jo) {
    for (int i = tri[j - 1].start_idx; i <= tri[j - 1].end_idx; i++) B.gp[i].sn += jo;
  };
  for (int j = 1; j <= 5; j++) shift(j, (j - 1) * off * 2 + 1);
  for (int j = 16; j <= 20; j++) shift(j, (j - 16) * off * 2 + 1);
  for (int j = 12; j <= 15; j++) shift(j, (j - 12) * off * 2 + off + 1);
  shift(11, 4 * off * 2 + off + 1);
  for (int j = 6; j <= 10; j++) shift(j, (j - 6) * off * 2 + off + 1);
  // gather (seqnum -> lat/lon in degrees), icos.F90:97-118 and GridGen.F90:75-87
  std::vector<double> glat(nip + 1), glon(nip + 1);
  std::vector<char> seen(nip + 1, 0);
  auto put = [&](int seq, double la, double lo) {
    if (seq < 1 || seq > nip) return false;
    glat[seq] = la;
    glon[seq] = lo;
    seen[seq] = 1;
    return true;
  };
  if (!put(1, top[0].ll[0], top[0].ll[1]) || !put(nip, top[6].ll[0], top[6].ll[1])) return FIMHIP_EINVAL;
  for (int j = 0; j < 20; j++)
    for (int i = tri[j].start_idx; i <= tri[j].end_idx; i++)
      if (!put(B.gp[i].sn + 1, B.gp[i].ll[0], B.gp[i].ll[1])) return FIMHIP_EINVAL;
  for (int i = 1; i <= nip; i++)
    if (!seen[i]) return FIMHIP_EINVAL;

  // ---- GridGen.F90: radians, neighbours, Voronoi corners ----------------
  const double d2r = (double)(4.0f * atanf(1.0f) / 180.0f);
  const double pi2 = (double)(8.0f * atanf(1.0f));
  const double eps = (double)1.0e-4f;
  std::vector<double> rlat(nip + 1), rlon(nip + 1);
  for (int i = 1; i <= nip; i++) {
    double lo = glon[i];
    if (lo < 0) lo = lo + 360.0;
    rlat[i] = glat[i] * d2r;
    rlon[i] = lo * d2r;
  }
  const int diam_sz = (nip - 2) / 10;
  const int edge_ln = (int)(sqrtf((float)diam_sz) + 0.5f);
  std::vector<int> prox((size_t)nip * 6), nprox(nip);
  std::vector<double> edge((size_t)nip * 24, 0.);   // icos_edge(6,2,2,nip) -> [(i*2+i3)*2+i2)*6+j]
  int nb[6] = {0, 0, 0, 0, 0, 0};
  double v[6][3];
  memset(v, 0, sizeof v);
  for (int i = 1; i <= nip; i++) {
    int np = 6;
    int sgp = special_points(i, nb, diam_sz, edge_ln);
    if (sgp == 1) np = 5;
    if (sgp == 0) pattern_neighbours(i, nb, diam_sz, edge_ln);
    double xyz[7][3];
    xyz[0][0] = cos(rlat[i]) * cos(rlon[i]);
    xyz[0][1] = cos(rlat[i]) * sin(rlon[i]);
    xyz[0][2] = sin(rlat[i]);
    int jmax = nb[5] == -1 ? 5 : 6;
    for (int j = 1; j <= jmax; j++) {
      int q = nb[j - 1];
      xyz[j][0] = cos(rlat[q]) * cos(rlon[q]);
      xyz[j][1] = cos(rlat[q]) * sin(rlon[q]);
      xyz[j][2] = sin(rlat[q]);
    }
    nprox[i - 1] = np;
    for (int j = 0; j < 6; j++) prox[(size_t)(i - 1) * 6 + j] = nb[j];
    for (int j = 1; j <= jmax; j++) {
      if (nb[j - 1] > 0) {
        int k = j + 1;
        if (k > jmax) k = k - jmax;
        double v1[3], v2[3];
        for (int c = 0; c < 3; c++) { v1[c] = xyz[j][c] - xyz[0][c]; v2[c] = xyz[k][c] - xyz[0][c]; }
        v[j - 1][0] = v1[1] * v2[2] - v1[2] * v2[1];
        v[j - 1][1] = v1[2] * v2[0] - v1[0] * v2[2];
        v[j - 1][2] = v1[0] * v2[1] - v1[1] * v2[0];
        double d = sqrt(dot3(v[j - 1], v[j - 1]));
        for (int c = 0; c < 3; c++) v[j - 1][c] = v[j - 1][c] / d;
      }
    }
    auto edge_ll = [&](const double *vv, double *la, double *lo) {
      *la = asin(vv[2]);
      if (fabs(*la) > pi2 / 4.0 - eps) {
        *lo = 0.0;
      } else {
        double q = vv[0] / cos(*la);
        if (q >= 1.0) *lo = 0.0;
        else if (q <= -1.0) *lo = pi2 / 2.0;
        else {
          *lo = acos(vv[0] / cos(*la));
          if (vv[1] < 0.0) *lo = pi2 - *lo;
        }
      }
    };
    for (int j = 1; j <= jmax; j++) {
      if (nb[j - 1] > 0) {
        int k = j + 1;
        if (k > jmax) k = k - jmax;
        double *e = &edge[(size_t)(i - 1) * 24];
        edge_ll(v[j - 1], &e[(0 * 2 + 0) * 6 + j - 1], &e[(1 * 2 + 0) * 6 + j - 1]);
        edge_ll(v[k - 1], &e[(0 * 2 + 1) * 6 + j - 1], &e[(1 * 2 + 1) * 6 + j - 1]);
      }
    }
  }

  // ---- getlvl.F90: FP32 metrics ------------------------------------------
  const float pi_g = (float)3.1415926535897;
  const float ae = 6371220.f;
  std::vector<float> lat(nip), lon(nip);
  for (int i = 0; i < nip; i++) {
    lat[i] = (float)rlat[i + 1];
    lon[i] = (float)rlon[i + 1];
    if (lon[i] < 0.f) lon[i] = lon[i] + 2.f * pi_g;
  }
  // conr_ll(isn, ipt, ixy, ipn) with the edge-index shift of getlvl.F90:129-145
  std::vector<float> conr((size_t)nip * 24);   // [((ipn*2+ixy)*2+ipt)*6+isn]
  auto CONR = [&](int ipn, int isn, int ipt, int ixy) -> float & {
    return conr[(((size_t)ipn * 2 + ixy) * 2 + ipt) * 6 + isn];
  };
  for (int ipn = 0; ipn < nip; ipn++) {
    const double *e = &edge[(size_t)ipn * 24];
    float tmp[6][2];
    int np = nprox[ipn];
    for (int isn = 0; isn < np; isn++)
      for (int ixy = 0; ixy < 2; ixy++) tmp[isn][ixy] = (float)e[(ixy * 2 + 0) * 6 + isn];
    for (int isn = 0; isn < 6; isn++)
      for (int ipt = 0; ipt < 2; ipt++)
        for (int ixy = 0; ixy < 2; ixy++) CONR(ipn, isn, ipt, ixy) = (float)e[(ixy * 2 + ipt) * 6 + isn];
    for (int isn = 0; isn < np; isn++) {
      int ism = isn == 0 ? np - 1 : isn - 1;
      for (int ixy = 0; ixy < 2; ixy++) {
        CONR(ipn, isn, 0, ixy) = tmp[ism][ixy];
        CONR(ipn, isn, 1, ixy) = tmp[isn][ixy];
      }
    }
  }
  // proxs (getlvl.F90:146-157)
  std::vector<int> proxs((size_t)nip * 6, -99);
  for (int ipn = 0; ipn < nip; ipn++)
    for (int isn = 0; isn < nprox[ipn]; isn++) {
      int ip = prox[(size_t)ipn * 6 + isn] - 1;
      for (int j = 0; j < nprox[ip]; j++)
        if (prox[(size_t)ip * 6 + j] == ipn + 1) { proxs[(size_t)ipn * 6 + isn] = j + 1; break; }
    }
  std::vector<float> lle((size_t)nip * 12, 0.f);   // lle(isn, 1:2, ipn) -> [(ipn*2+c)*6+isn]
  for (int ipn = 0; ipn < nip; ipn++)
    for (int isn = 0; isn < nprox[ipn]; isn++) {
      float p1[2] = {CONR(ipn, isn, 0, 0), CONR(ipn, isn, 0, 1)};
      float p2[2] = {CONR(ipn, isn, 1, 0), CONR(ipn, isn, 1, 1)};
      float pm[2];
      middle_r(p1, p2, pm);
      lle[((size_t)ipn * 2 + 0) * 6 + isn] = pm[0];
      lle[((size_t)ipn * 2 + 1) * 6 + isn] = pm[1];
    }
  std::vector<float> area(nip), cs((size_t)nip * 24, 0.f), sn((size_t)nip * 24, 0.f);
  std::vector<float> svc((size_t)nip * 12, 0.f), sve((size_t)nip * 12, 0.f), sideln((size_t)nip * 6, 0.f);
  std::vector<float> rprox((size_t)nip * 6, 0.f), corner((size_t)nip * 12, 0.f), hf((size_t)nip * 36, 0.f);
  std::vector<float> pxy((size_t)nip * 30, 0.f);   // pxy(5,6,nip) -> [(ipn*6+isn)*5+b]
  for (int ipn = 0; ipn < nip; ipn++) {
    const int np = nprox[ipn];
    float conr_xy[6][2][2];
    for (int isn = 0; isn < np; isn++) {
      int q = prox[(size_t)ipn * 6 + isn] - 1;
      float px, py;
      ll2xy(lon[ipn], lat[ipn], lon[q], lat[q], &px, &py);
      rprox[(size_t)ipn * 6 + isn] = 1.f / (ae * sqrtf(px * px + py * py));
      for (int ipt = 0; ipt < 2; ipt++)
        ll2xy(lon[ipn], lat[ipn], CONR(ipn, isn, ipt, 1), CONR(ipn, isn, ipt, 0), &conr_xy[isn][ipt][0],
              &conr_xy[isn][ipt][1]);
      const float l1 = lle[((size_t)ipn * 2 + 0) * 6 + isn], l2 = lle[((size_t)ipn * 2 + 1) * 6 + isn];
      float map = 2.f / (1.f + sinf(l1) * sinf(lat[ipn]) + cosf(l1) * cosf(lat[ipn]) * cosf(l2 - lon[ipn]));
      for (int ixy = 0; ixy < 2; ixy++)
        svc[((size_t)ipn * 6 + isn) * 2 + ixy] = ae * (conr_xy[isn][1][ixy] - conr_xy[isn][0][ixy]) * map;
      float xxp, yyp, xxm, yym;
      ll2xy(l2, l1, CONR(ipn, isn, 1, 1), CONR(ipn, isn, 1, 0), &xxp, &yyp);
      ll2xy(l2, l1, CONR(ipn, isn, 0, 1), CONR(ipn, isn, 0, 0), &xxm, &yym);
      float e1 = ae * (xxp - xxm), e2 = ae * (yyp - yym);
      sve[((size_t)ipn * 6 + isn) * 2 + 0] = e1;
      sve[((size_t)ipn * 6 + isn) * 2 + 1] = e2;
      sideln[(size_t)ipn * 6 + isn] = sqrtf(e1 * e1 + e2 * e2);
    }
    float ar = 0.f;
    for (int isn = 0; isn < np; isn++) {
      float xx = ae * .5f * (conr_xy[isn][1][0] + conr_xy[isn][0][0]);
      float yy = ae * .5f * (conr_xy[isn][1][1] + conr_xy[isn][0][1]);
      ar = ar + .5f * (xx * svc[((size_t)ipn * 6 + isn) * 2 + 1] - yy * svc[((size_t)ipn * 6 + isn) * 2 + 0]);
    }
    area[ipn] = ar;
    for (int ixy = 0; ixy < 2; ixy++)
      for (int isn = 0; isn < np; isn++) corner[((size_t)ipn * 2 + ixy) * 6 + isn] = ae * conr_xy[isn][0][ixy];
    // moments of the cell polygon and pxy (getlvl.F90:219-248)
    float ap[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    for (int isn = 0; isn < np; isn++) {
      float xim1 = conr_xy[isn][0][0], xi = conr_xy[isn][1][0];
      float yim1 = conr_xy[isn][0][1], yi = conr_xy[isn][1][1];
      float xx = ae * .5f * (xim1 + xi);
      float yy = ae * .5f * (yim1 + yi);
      float tri = .5f * (xx * svc[((size_t)ipn * 6 + isn) * 2 + 1] - yy * svc[((size_t)ipn * 6 + isn) * 2 + 0]) / area[ipn];
      ap[0] = ap[0] + tri / 3.f * (xim1 + xi);
      ap[1] = ap[1] + tri / 3.f * (yim1 + yi);
      ap[2] = ap[2] + tri / 6.f * (xim1 * xim1 + xim1 * xi + xi * xi);
      ap[3] = ap[3] + tri / 12.f * (2.f * xim1 * yim1 + 2.f * xi * yi + xim1 * yi + xi * yim1);
      ap[4] = ap[4] + tri / 6.f * (yim1 * yim1 + yim1 * yi + yi * yi);
    }
    for (int isn = 0; isn < np; isn++) {
      float xim1 = conr_xy[isn][0][0], xi = conr_xy[isn][1][0];
      float yim1 = conr_xy[isn][0][1], yi = conr_xy[isn][1][1];
      float *p = &pxy[((size_t)ipn * 6 + isn) * 5];
      p[0] = .5f * (xim1 + xi) - ap[0];
      p[1] = .5f * (yim1 + yi) - ap[1];
      p[2] = 1.f / 3.f * (xim1 * xim1 + xi * xi + xim1 * xi) - ap[2];
      p[3] = 1.f / 6.f * (2.f * xim1 * yim1 + 2.f * xi * yi + xim1 * yi + xi * yim1) - ap[3];
      p[4] = 1.f / 3.f * (yim1 * yim1 + yi * yi + yim1 * yi) - ap[4];
    }
    // hfcoef (getlvl.F90:250-269)
    float loc[6 * 2], pinv[5 * 6];
    for (int isn = 0; isn < np; isn++) {
      int q = prox[(size_t)ipn * 6 + isn] - 1;
      ll2xy(lon[ipn], lat[ipn], lon[q], lat[q], &loc[2 * isn], &loc[2 * isn + 1]);
    }
    if (fimhip_ls_pseudo_inverse(loc, np, pinv) != FIMHIP_OK) return FIMHIP_EINVAL;
    for (int isn = 0; isn < np; isn++) {
      float *h = &hf[((size_t)ipn * 6 + isn) * 6];
      for (int ico = 0; ico < 6; ico++) h[ico] = 0.f;
      for (int ico = 0; ico < np; ico++)
        for (int ib = 0; ib < 5; ib++) h[ico] = h[ico] + pxy[((size_t)ipn * 6 + isn) * 5 + ib] * pinv[ib + ico * 5];
    }
    // cs, sn (getlvl.F90:273-294)
    for (int isn = 0; isn < np; isn++) {
      float xltc = lle[((size_t)ipn * 2 + 0) * 6 + isn], xlnc = lle[((size_t)ipn * 2 + 1) * 6 + isn];
      int ip1 = (isn + 1) % np, im1 = isn == 0 ? np - 1 : isn - 1;
      int pr[4] = {ipn, prox[(size_t)ipn * 6 + isn] - 1, prox[(size_t)ipn * 6 + im1] - 1,
                   prox[(size_t)ipn * 6 + ip1] - 1};
      for (int ipt = 0; ipt < 4; ipt++) {
        float elt = lat[pr[ipt]], eln = lon[pr[ipt]];
        float rf = 1.0f / (1.0f + sinf(xltc) * sinf(elt) + cosf(xltc) * cosf(elt) * cosf(eln - xlnc));
        cs[((size_t)ipn * 6 + isn) * 4 + ipt] = rf * (cosf(xltc) * cosf(elt) + (1.0f + sinf(xltc) * sinf(elt)) * cosf(eln - xlnc));
        sn[((size_t)ipn * 6 + isn) * 4 + ipt] = -rf * sinf(eln - xlnc) * (sinf(xltc) + sinf(elt));
      }
    }
  }
  // ---- outputs (Fortran layout, as glvl.dat) --------------------------------
  for (int i = 0; i < nip; i++) {
    if (lat_out) lat_out[i] = lat[i];
    if (lon_out) lon_out[i] = lon[i];
    if (nprox_out) nprox_out[i] = nprox[i];
    if (area_out) area_out[i] = area[i];
    if (inv_perm_out) inv_perm_out[i] = i + 1;
    for (int j = 0; j < 6; j++) {
      if (prox_out) prox_out[i * 6 + j] = prox[(size_t)i * 6 + j];
      if (proxs_out) proxs_out[i * 6 + j] = proxs[(size_t)i * 6 + j];
      if (sideln_out) sideln_out[i * 6 + j] = sideln[(size_t)i * 6 + j];
      if (rprox_out) rprox_out[i * 6 + j] = rprox[(size_t)i * 6 + j];
    }
  }
  if (cs_out) memcpy(cs_out, cs.data(), cs.size() * sizeof(float));
  if (sn_out) memcpy(sn_out, sn.data(), sn.size() * sizeof(float));
  if (svc_out) memcpy(svc_out, svc.data(), svc.size() * sizeof(float));
  if (sve_out) memcpy(sve_out, sve.data(), sve.size() * sizeof(float));
  if (corner_out) memcpy(corner_out, corner.data(), corner.size() * sizeof(float));
  if (hfcoef_out) memcpy(hfcoef_out, hf.data(), hf.size() * sizeof(float));
  if (edge_out)
    for (size_t i = 0; i < edge.size(); i++) edge_out[i] = (float)edge[i];
  return FIMHIP_OK;
}

extern "C" int fimhip_icos_grid(int glvl, float *lat_out, float *lon_out, int *nprox_out, int *prox_out,
                                int *proxs_out, float *area_out, float *cs_out, float *sn_out,
                                float *svc_out, float *sve_out, float *sideln_out, float *rprox_out,
                                float *corner_out, float *hfcoef_out, int *inv_perm_out) {
  return icos_grid_impl(glvl, lat_out, lon_out, nprox_out, prox_out, proxs_out, area_out, cs_out, sn_out, svc_out,
                        sve_out, sideln_out, rprox_out, corner_out, hfcoef_out, inv_perm_out, nullptr);
}

// The ginfo output (GridGen.F90:437-497, single precision): cell centres in
// radians, the pattern neighbours (prox(6,nip), slot 6 of a pentagon as the
// pattern leaves it), nprox, the Voronoi corner of each edge end
// (icos_edge(6,2,2,nip): edge, end, lat/lon) and inv_perm.
extern "C" int fimhip_icos_grid_info(int glvl, float *lat_out, float *lon_out, int *prox_out, int *nprox_out,
                                     float *edge_out, int *inv_perm_out) {
  return icos_grid_impl(glvl, lat_out, lon_out, nprox_out, prox_out, nullptr, nullptr, nullptr, nullptr, nullptr,
                        nullptr, nullptr, nullptr, nullptr, nullptr, inv_perm_out, edge_out);
}
