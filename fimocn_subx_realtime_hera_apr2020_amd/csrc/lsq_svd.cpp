// The least-squares pseudo-inverse behind the edge reconstruction
// coefficients hfcoef (getlvl.F90:250-269 calls LSfitting.F90:73-130,
// pseudoInverse): the singular value decomposition A = U S V^T of the
// nobs x 5 observation matrix (nobs = 5 or 6) in single precision, by the
// algorithm of the LAPACK release the reference vendors in prep/lsfit
// (svd.f, la_aux.f, blas.f), then inv = V S^-1 U^T.
//
// Bit-exact results need the same float operations in the same order, so
// each routine below follows the vendored routine it names step by step:
// Householder bidiagonalisation (sgebd2 with slarfg/slarf), the explicit
// orthogonal factors (sorgbr through sorg2r/sorgl2), and the implicit-shift
// bidiagonal QR iteration with its convergence tests, zero-shift sweeps and
// final sort (sbdsqr with slartg, slas2, slasv2, slasr).  Only the branches
// that sgesvd('A','A') takes for these sizes are restated: no workspace
// queries, no blocked code (the sizes are far below the block size), and no
// rescaling (the matrix entries are stereographic coordinates and their
// products, far inside the range where sgesvd would rescale; that case is
// rejected).  Matrices are column-major with 1-based indices, as in the
// routines followed.
#include <math.h>
#include <string.h>

#include "fimhip.h"

namespace {

const float EPS = ldexpf(1.f, -24);        // slamch('E'): EPSILON * 0.5 (la_aux.f:1673)
const float SFMIN = ldexpf(1.f, -126);     // slamch('S'): TINY (1/HUGE is smaller)

struct Mat {                               // A(i, j), 1-based, leading dimension ld
  float *p;
  int ld;
  float &operator()(int i, int j) const { return p[(i - 1) + (long)(j - 1) * ld]; }
  Mat sub(int i, int j) const { return Mat{&(*this)(i, j), ld}; }
};
struct Vec {                               // x(i), 1-based, stride inc
  float *p;
  int inc;
  float &operator()(int i) const { return p[(long)(i - 1) * inc]; }
};

float fsign(float a, float b) { return copysignf(fabsf(a), b); }   // Fortran SIGN(a, b)
float fmaxf2(float a, float b) { return a > b ? a : b; }
float fminf2(float a, float b) { return a < b ? a : b; }

// snrm2 (blas.f:22794-22860): scaled sum of squares
float snrm2(int n, Vec x) {
  if (n < 1) return 0.f;
  if (n == 1) return fabsf(x(1));
  float scale = 0.f, ssq = 1.f;
  for (int i = 1; i <= n; i++) {
    if (x(i) != 0.f) {
      const float absxi = fabsf(x(i));
      if (scale < absxi) {
        const float r = scale / absxi;
        ssq = 1.f + ssq * (r * r);
        scale = absxi;
      } else {
        const float r = absxi / scale;
        ssq = ssq + r * r;
      }
    }
  }
  return scale * sqrtf(ssq);
}

// slapy2 (svd.f:5903): sqrt(x^2 + y^2) without overflow
float slapy2(float x, float y) {
  const float xa = fabsf(x), ya = fabsf(y);
  const float w = fmaxf2(xa, ya), z = fminf2(xa, ya);
  if (z == 0.f) return w;
  const float r = z / w;
  return w * sqrtf(1.f + r * r);
}

void sscal(int n, float a, Vec x) {
  for (int i = 1; i <= n; i++) x(i) = a * x(i);
}

// slarfg (svd.f:6777): elementary reflector H with H (alpha, x) = (beta, 0)
void slarfg(int n, float &alpha, Vec x, float &tau) {
  if (n <= 1) {
    tau = 0.f;
    return;
  }
  float xnorm = snrm2(n - 1, x);
  if (xnorm == 0.f) {
    tau = 0.f;
    return;
  }
  float beta = -fsign(slapy2(alpha, xnorm), alpha);
  const float safmin = SFMIN / EPS;
  int knt = 0;
  if (fabsf(beta) < safmin) {
    const float rsafmn = 1.f / safmin;
    do {
      knt++;
      sscal(n - 1, rsafmn, x);
      beta = beta * rsafmn;
      alpha = alpha * rsafmn;
    } while (fabsf(beta) < safmin);
    xnorm = snrm2(n - 1, x);
    beta = -fsign(slapy2(alpha, xnorm), alpha);
  }
  tau = (beta - alpha) / beta;
  sscal(n - 1, 1.f / (alpha - beta), x);
  for (int j = 1; j <= knt; j++) beta = beta * safmin;
  alpha = beta;
}

// ilaslc / ilaslr (la_aux.f:925, 1043): last non-zero column / row
int ilaslc(int m, int n, Mat a) {
  if (n == 0) return n;
  if (a(1, n) != 0.f || a(m, n) != 0.f) return n;
  for (int j = n; j >= 1; j--)
    for (int i = 1; i <= m; i++)
      if (a(i, j) != 0.f) return j;
  return 0;
}
int ilaslr(int m, int n, Mat a) {
  if (m == 0) return m;
  if (a(m, 1) != 0.f || a(m, n) != 0.f) return m;
  int r = 0;
  for (int j = 1; j <= n; j++) {
    int i = m;
    while (i >= 1 && a(i > 1 ? i : 1, j) == 0.f) i--;
    if (i > r) r = i;
  }
  return r;
}

// sgemv with alpha = 1, beta = 0 (blas.f:22339): y = A^T x or y = A x
void sgemv_t(int m, int n, Mat a, Vec x, Vec y) {
  for (int j = 1; j <= n; j++) y(j) = 0.f;
  for (int j = 1; j <= n; j++) {
    float temp = 0.f;
    for (int i = 1; i <= m; i++) temp = temp + a(i, j) * x(i);
    y(j) = y(j) + 1.f * temp;
  }
}
void sgemv_n(int m, int n, Mat a, Vec x, Vec y) {
  for (int i = 1; i <= m; i++) y(i) = 0.f;
  for (int j = 1; j <= n; j++) {
    const float temp = 1.f * x(j);
    for (int i = 1; i <= m; i++) y(i) = y(i) + temp * a(i, j);
  }
}
// sger (blas.f:22643): A = A + alpha x y^T
void sger(int m, int n, float alpha, Vec x, Vec y, Mat a) {
  for (int j = 1; j <= n; j++) {
    if (y(j) != 0.f) {
      const float temp = alpha * y(j);
      for (int i = 1; i <= m; i++) a(i, j) = a(i, j) + x(i) * temp;
    }
  }
}

// slarf (svd.f:6624): apply H = I - tau v v^T from the left or the right
void slarf(bool left, int m, int n, Vec v, float tau, Mat c, float *work) {
  int lastv = 0, lastc = 0;
  if (tau != 0.f) {
    lastv = left ? m : n;
    while (lastv > 0 && v(lastv) == 0.f) lastv--;
    lastc = left ? ilaslc(lastv, n, c) : ilaslr(m, lastv, c);
  }
  if (lastv <= 0) return;
  Vec w{work, 1};
  if (left) {
    sgemv_t(lastv, lastc, c, v, w);
    sger(lastv, lastc, -tau, v, w, c);
  } else {
    sgemv_n(lastc, lastv, c, v, w);
    sger(lastc, lastv, -tau, w, v, c);
  }
}

// sgebd2, m >= n (svd.f:751): A = Q B P^T, B upper bidiagonal (d, e)
void sgebd2(int m, int n, Mat a, float *d, float *e, float *tauq, float *taup, float *work) {
  for (int i = 1; i <= n; i++) {
    slarfg(m - i + 1, a(i, i), Vec{&a(i + 1 < m ? i + 1 : m, i), 1}, tauq[i - 1]);
    d[i - 1] = a(i, i);
    a(i, i) = 1.f;
    if (i < n) slarf(true, m - i + 1, n - i, Vec{&a(i, i), 1}, tauq[i - 1], a.sub(i, i + 1), work);
    a(i, i) = d[i - 1];
    if (i < n) {
      slarfg(n - i, a(i, i + 1), Vec{&a(i, i + 2 < n ? i + 2 : n), a.ld}, taup[i - 1]);
      e[i - 1] = a(i, i + 1);
      a(i, i + 1) = 1.f;
      slarf(false, m - i, n - i, Vec{&a(i, i + 1), a.ld}, taup[i - 1], a.sub(i + 1, i + 1), work);
      a(i, i + 1) = e[i - 1];
    } else {
      taup[i - 1] = 0.f;
    }
  }
}

// sorg2r (svd.f:10452): the m x n matrix Q of k reflectors stored in A
void sorg2r(int m, int n, int k, Mat a, const float *tau, float *work) {
  for (int j = k + 1; j <= n; j++) {
    for (int l = 1; l <= m; l++) a(l, j) = 0.f;
    a(j, j) = 1.f;
  }
  for (int i = k; i >= 1; i--) {
    if (i < n) {
      a(i, i) = 1.f;
      slarf(true, m - i + 1, n - i, Vec{&a(i, i), 1}, tau[i - 1], a.sub(i, i + 1), work);
    }
    if (i < m) sscal(m - i, -tau[i - 1], Vec{&a(i + 1, i), 1});
    a(i, i) = 1.f - tau[i - 1];
    for (int l = 1; l <= i - 1; l++) a(l, i) = 0.f;
  }
}

// sorgl2 (svd.f:10827): the m x n matrix Q^T of k row reflectors stored in A
void sorgl2(int m, int n, int k, Mat a, const float *tau, float *work) {
  if (m <= 0) return;
  if (k < m) {
    for (int j = 1; j <= n; j++) {
      for (int l = k + 1; l <= m; l++) a(l, j) = 0.f;
      if (j > k && j <= m) a(j, j) = 1.f;
    }
  }
  for (int i = k; i >= 1; i--) {
    if (i < n) {
      if (i < m) {
        a(i, i) = 1.f;
        slarf(false, m - i, n - i + 1, Vec{&a(i, i), a.ld}, tau[i - 1], a.sub(i + 1, i), work);
      }
      sscal(n - i, -tau[i - 1], Vec{&a(i, i + 1), a.ld});
    }
    a(i, i) = 1.f - tau[i - 1];
    for (int l = 1; l <= i - 1; l++) a(i, l) = 0.f;
  }
}

// sorgbr('P', n, n, n) (svd.f:10582, k >= n): shift the row reflectors one
// row down and form P^T(2:n, 2:n) (sorglq at these sizes is sorgl2)
void sorgbr_p(int n, Mat a, const float *taup, float *work) {
  a(1, 1) = 1.f;
  for (int i = 2; i <= n; i++) a(i, 1) = 0.f;
  for (int j = 2; j <= n; j++) {
    for (int i = j - 1; i >= 2; i--) a(i, j) = a(i - 1, j);
    a(1, j) = 0.f;
  }
  if (n > 1) sorgl2(n - 1, n - 1, n - 1, a.sub(2, 2), taup, work);
}

// slartg (svd.f:7163): plane rotation with cs f + sn g = r, -sn f + cs g = 0
void slartg(float f, float g, float &cs, float &sn, float &r) {
  const float safmin = SFMIN, eps = EPS;
  const int ex = (int)(logf(safmin / eps) / logf(2.f) / 2.f);
  float safmn2 = 1.f;                      // 2 ** ex (ex < 0)
  for (int i = 0; i < -ex; i++) safmn2 = safmn2 / 2.f;
  const float safmx2 = 1.f / safmn2;
  if (g == 0.f) {
    cs = 1.f;
    sn = 0.f;
    r = f;
  } else if (f == 0.f) {
    cs = 0.f;
    sn = 1.f;
    r = g;
  } else {
    float f1 = f, g1 = g;
    float scale = fmaxf2(fabsf(f1), fabsf(g1));
    if (scale >= safmx2) {
      int count = 0;
      do {
        count++;
        f1 = f1 * safmn2;
        g1 = g1 * safmn2;
        scale = fmaxf2(fabsf(f1), fabsf(g1));
      } while (scale >= safmx2);
      r = sqrtf(f1 * f1 + g1 * g1);
      cs = f1 / r;
      sn = g1 / r;
      for (int i = 1; i <= count; i++) r = r * safmx2;
    } else if (scale <= safmn2) {
      int count = 0;
      do {
        count++;
        f1 = f1 * safmx2;
        g1 = g1 * safmx2;
        scale = fmaxf2(fabsf(f1), fabsf(g1));
      } while (scale <= safmn2);
      r = sqrtf(f1 * f1 + g1 * g1);
      cs = f1 / r;
      sn = g1 / r;
      for (int i = 1; i <= count; i++) r = r * safmn2;
    } else {
      r = sqrtf(f1 * f1 + g1 * g1);
      cs = f1 / r;
      sn = g1 / r;
    }
    if (fabsf(f) > fabsf(g) && cs < 0.f) {
      cs = -cs;
      sn = -sn;
      r = -r;
    }
  }
}

// slas2 (svd.f:7309): singular values of [[f, g], [0, h]]
void slas2(float f, float g, float h, float &ssmin, float &ssmax) {
  const float fa = fabsf(f), ga = fabsf(g), ha = fabsf(h);
  const float fhmn = fminf2(fa, ha), fhmx = fmaxf2(fa, ha);
  if (fhmn == 0.f) {
    ssmin = 0.f;
    if (fhmx == 0.f) {
      ssmax = ga;
    } else {
      const float q = fminf2(fhmx, ga) / fmaxf2(fhmx, ga);
      ssmax = fmaxf2(fhmx, ga) * sqrtf(1.f + q * q);
    }
  } else if (ga < fhmx) {
    const float as = 1.f + fhmn / fhmx, at = (fhmx - fhmn) / fhmx, q = ga / fhmx, au = q * q;
    const float c = 2.f / (sqrtf(as * as + au) + sqrtf(at * at + au));
    ssmin = fhmn * c;
    ssmax = fhmx / c;
  } else {
    const float au = fhmx / ga;
    if (au == 0.f) {
      ssmin = (fhmn * fhmx) / ga;
      ssmax = ga;
    } else {
      const float as = 1.f + fhmn / fhmx, at = (fhmx - fhmn) / fhmx;
      const float p = as * au, q = at * au;
      const float c = 1.f / (sqrtf(1.f + p * p) + sqrtf(1.f + q * q));
      ssmin = (fhmn * c) * au;
      ssmin = ssmin + ssmin;
      ssmax = ga / (c + c);
    }
  }
}

// slasv2 (svd.f:10202): SVD of [[f, g], [0, h]] with the rotations
void slasv2(float f, float g, float h, float &ssmin, float &ssmax, float &snr, float &csr, float &snl,
            float &csl) {
  float ft = f, fa = fabsf(ft), ht = h, ha = fabsf(h);
  int pmax = 1;
  const bool swap = ha > fa;
  if (swap) {
    pmax = 3;
    float t = ft;
    ft = ht;
    ht = t;
    t = fa;
    fa = ha;
    ha = t;
  }
  const float gt = g, ga = fabsf(gt);
  float clt, crt, slt, srt;
  if (ga == 0.f) {
    ssmin = ha;
    ssmax = fa;
    clt = 1.f;
    crt = 1.f;
    slt = 0.f;
    srt = 0.f;
  } else {
    bool gasmal = true;
    if (ga > fa) {
      pmax = 2;
      if ((fa / ga) < EPS) {
        gasmal = false;
        ssmax = ga;
        if (ha > 1.f) ssmin = fa / (ga / ha);
        else ssmin = (fa / ga) * ha;
        clt = 1.f;
        slt = ht / gt;
        srt = 1.f;
        crt = ft / gt;
      }
    }
    if (gasmal) {
      const float d = fa - ha;
      float l = d == fa ? 1.f : d / fa;
      const float m = gt / ft;
      float t = 2.f - l;
      const float mm = m * m, tt = t * t;
      const float s = sqrtf(tt + mm);
      const float r = l == 0.f ? fabsf(m) : sqrtf(l * l + mm);
      const float a = .5f * (s + r);
      ssmin = ha / a;
      ssmax = fa * a;
      if (mm == 0.f) {
        if (l == 0.f) t = fsign(2.f, ft) * fsign(1.f, gt);
        else t = gt / fsign(d, ft) + m / t;
      } else {
        t = (m / (s + t) + m / (r + l)) * (1.f + a);
      }
      l = sqrtf(t * t + 4.f);
      crt = 2.f / l;
      srt = t / l;
      clt = (crt + srt * m) / a;
      slt = (ht / ft) * srt / a;
    }
  }
  if (swap) {
    csl = srt;
    snl = crt;
    csr = slt;
    snr = clt;
  } else {
    csl = clt;
    snl = slt;
    csr = crt;
    snr = srt;
  }
  float tsign = 1.f;
  if (pmax == 1) tsign = fsign(1.f, csr) * fsign(1.f, csl) * fsign(1.f, f);
  if (pmax == 2) tsign = fsign(1.f, snr) * fsign(1.f, csl) * fsign(1.f, g);
  if (pmax == 3) tsign = fsign(1.f, snr) * fsign(1.f, snl) * fsign(1.f, h);
  ssmax = fsign(ssmax, tsign);
  ssmin = fsign(ssmin, tsign * fsign(1.f, f) * fsign(1.f, h));
}

// slasr with pivot 'V' (svd.f:9507): a sequence of plane rotations applied
// from the left (to rows j, j+1) or the right (to columns j, j+1), forward
// or backward
void slasr_v(bool left, bool forward, int m, int n, const float *c, const float *s, Mat a) {
  if (m == 0 || n == 0) return;
  const int nr = (left ? m : n) - 1;
  for (int q = 1; q <= nr; q++) {
    const int j = forward ? q : nr + 1 - q;
    const float ct = c[j - 1], st = s[j - 1];
    if (ct == 1.f && st == 0.f) continue;
    if (left) {
      for (int i = 1; i <= n; i++) {
        const float temp = a(j + 1, i);
        a(j + 1, i) = ct * temp - st * a(j, i);
        a(j, i) = st * temp + ct * a(j, i);
      }
    } else {
      for (int i = 1; i <= m; i++) {
        const float temp = a(i, j + 1);
        a(i, j + 1) = ct * temp - st * a(i, j);
        a(i, j) = st * temp + ct * a(i, j);
      }
    }
  }
}

// srot (blas.f:22903)
void srot(int n, Vec x, Vec y, float c, float s) {
  for (int i = 1; i <= n; i++) {
    const float t = c * x(i) + s * y(i);
    y(i) = c * y(i) - s * x(i);
    x(i) = t;
  }
}
void sswap(int n, Vec x, Vec y) {
  for (int i = 1; i <= n; i++) {
    const float t = x(i);
    x(i) = y(i);
    y(i) = t;
  }
}

// sbdsqr('U', n, ncvt, nru, 0) (svd.f:1-750): the SVD of the upper
// bidiagonal (d, e), the right rotations applied to the rows of vt
// (n x ncvt) and the left ones to the columns of u (nru x n).  Returns the
// number of off-diagonals that failed to converge (0 on success).
int sbdsqr_u(int n, int ncvt, int nru, float *d, float *e, Mat vt, Mat u, float *work) {
  if (n == 0) return 0;
  auto D = [&](int i) -> float & { return d[i - 1]; };
  auto E = [&](int i) -> float & { return e[i - 1]; };
  auto W = [&](int i) -> float & { return work[i - 1]; };
  if (n > 1) {
    const int nm1 = n - 1, nm12 = nm1 + nm1, nm13 = nm12 + nm1;
    int idir = 0;
    const float eps = EPS, unfl = SFMIN;
    const float tolmul = fmaxf2(10.f, fminf2(100.f, powf(eps, -0.125f)));
    const float tol = tolmul * eps;
    float smax = 0.f;
    for (int i = 1; i <= n; i++) smax = fmaxf2(smax, fabsf(D(i)));
    for (int i = 1; i <= n - 1; i++) smax = fmaxf2(smax, fabsf(E(i)));
    float sminl = 0.f, thresh;
    {
      float sminoa = fabsf(D(1));
      if (sminoa != 0.f) {
        float mu = sminoa;
        for (int i = 2; i <= n; i++) {
          mu = fabsf(D(i)) * (mu / (mu + fabsf(E(i - 1))));
          sminoa = fminf2(sminoa, mu);
          if (sminoa == 0.f) break;
        }
      }
      sminoa = sminoa / sqrtf((float)n);
      thresh = fmaxf2(tol * sminoa, (float)(6 * n * n) * unfl);
    }
    const int maxit = 6 * n * n;
    int iter = 0, oldll = -1, oldm = -1, m = n;
    float smin = 0.f;
    for (;;) {
      if (m <= 1) break;
      if (iter > maxit) {
        int info = 0;
        for (int i = 1; i <= n - 1; i++)
          if (E(i) != 0.f) info++;
        return info;
      }
      // find the diagonal block to work on
      smax = fabsf(D(m));
      smin = smax;
      int ll = 0;
      bool split = false;
      for (int lll = 1; lll <= m - 1; lll++) {
        ll = m - lll;
        const float abss = fabsf(D(ll)), abse = fabsf(E(ll));
        if (abse <= thresh) {
          split = true;
          break;
        }
        smin = fminf2(smin, abss);
        smax = fmaxf2(fmaxf2(smax, abss), abse);
      }
      if (split) {
        E(ll) = 0.f;
        if (ll == m - 1) {
          m = m - 1;
          continue;
        }
      } else {
        ll = 0;
      }
      ll = ll + 1;
      if (ll == m - 1) {            // 2 x 2 block
        float sigmn, sigmx, sinr, cosr, sinl, cosl;
        slasv2(D(m - 1), E(m - 1), D(m), sigmn, sigmx, sinr, cosr, sinl, cosl);
        D(m - 1) = sigmx;
        E(m - 1) = 0.f;
        D(m) = sigmn;
        if (ncvt > 0) srot(ncvt, Vec{&vt(m - 1, 1), vt.ld}, Vec{&vt(m, 1), vt.ld}, cosr, sinr);
        if (nru > 0) srot(nru, Vec{&u(1, m - 1), 1}, Vec{&u(1, m), 1}, cosl, sinl);
        m = m - 2;
        continue;
      }
      if (ll > oldm || m < oldll) idir = fabsf(D(ll)) >= fabsf(D(m)) ? 1 : 2;
      // convergence tests
      bool conv = false;
      if (idir == 1) {
        if (fabsf(E(m - 1)) <= fabsf(tol) * fabsf(D(m))) {
          E(m - 1) = 0.f;
          continue;
        }
        float mu = fabsf(D(ll));
        sminl = mu;
        for (int lll = ll; lll <= m - 1; lll++) {
          if (fabsf(E(lll)) <= tol * mu) {
            E(lll) = 0.f;
            conv = true;
            break;
          }
          mu = fabsf(D(lll + 1)) * (mu / (mu + fabsf(E(lll))));
          sminl = fminf2(sminl, mu);
        }
      } else {
        if (fabsf(E(ll)) <= fabsf(tol) * fabsf(D(ll))) {
          E(ll) = 0.f;
          continue;
        }
        float mu = fabsf(D(m));
        sminl = mu;
        for (int lll = m - 1; lll >= ll; lll--) {
          if (fabsf(E(lll)) <= tol * mu) {
            E(lll) = 0.f;
            conv = true;
            break;
          }
          mu = fabsf(D(lll)) * (mu / (mu + fabsf(E(lll))));
          sminl = fminf2(sminl, mu);
        }
      }
      if (conv) continue;
      oldll = ll;
      oldm = m;
      // shift
      float shift, r;
      if ((float)n * tol * (sminl / smax) <= fmaxf2(eps, .01f * tol)) {
        shift = 0.f;
      } else {
        float sll;
        if (idir == 1) {
          sll = fabsf(D(ll));
          slas2(D(m - 1), E(m - 1), D(m), shift, r);
        } else {
          sll = fabsf(D(m));
          slas2(D(ll), E(ll), D(ll + 1), shift, r);
        }
        if (sll > 0.f) {
          const float q = shift / sll;
          if (q * q < eps) shift = 0.f;
        }
      }
      iter = iter + m - ll;
      if (shift == 0.f) {
        if (idir == 1) {               // zero-shift sweep, top to bottom
          float cs = 1.f, oldcs = 1.f, sn = 0.f, oldsn = 0.f;
          for (int i = ll; i <= m - 1; i++) {
            slartg(D(i) * cs, E(i), cs, sn, r);
            if (i > ll) E(i - 1) = oldsn * r;
            slartg(oldcs * r, D(i + 1) * sn, oldcs, oldsn, D(i));
            W(i - ll + 1) = cs;
            W(i - ll + 1 + nm1) = sn;
            W(i - ll + 1 + nm12) = oldcs;
            W(i - ll + 1 + nm13) = oldsn;
          }
          const float h = D(m) * cs;
          D(m) = h * oldcs;
          E(m - 1) = h * oldsn;
          if (ncvt > 0) slasr_v(true, true, m - ll + 1, ncvt, &W(1), &W(n), vt.sub(ll, 1));
          if (nru > 0) slasr_v(false, true, nru, m - ll + 1, &W(nm12 + 1), &W(nm13 + 1), u.sub(1, ll));
          if (fabsf(E(m - 1)) <= thresh) E(m - 1) = 0.f;
        } else {                       // zero-shift sweep, bottom to top
          float cs = 1.f, oldcs = 1.f, sn = 0.f, oldsn = 0.f;
          for (int i = m; i >= ll + 1; i--) {
            slartg(D(i) * cs, E(i - 1), cs, sn, r);
            if (i < m) E(i) = oldsn * r;
            slartg(oldcs * r, D(i - 1) * sn, oldcs, oldsn, D(i));
            W(i - ll) = cs;
            W(i - ll + nm1) = -sn;
            W(i - ll + nm12) = oldcs;
            W(i - ll + nm13) = -oldsn;
          }
          const float h = D(ll) * cs;
          D(ll) = h * oldcs;
          E(ll) = h * oldsn;
          if (ncvt > 0) slasr_v(true, false, m - ll + 1, ncvt, &W(nm12 + 1), &W(nm13 + 1), vt.sub(ll, 1));
          if (nru > 0) slasr_v(false, false, nru, m - ll + 1, &W(1), &W(n), u.sub(1, ll));
          if (fabsf(E(ll)) <= thresh) E(ll) = 0.f;
        }
      } else {
        if (idir == 1) {               // shifted QR sweep, top to bottom
          float f = (fabsf(D(ll)) - shift) * (fsign(1.f, D(ll)) + shift / D(ll));
          float g = E(ll);
          for (int i = ll; i <= m - 1; i++) {
            float cosr, sinr, cosl, sinl;
            slartg(f, g, cosr, sinr, r);
            if (i > ll) E(i - 1) = r;
            f = cosr * D(i) + sinr * E(i);
            E(i) = cosr * E(i) - sinr * D(i);
            g = sinr * D(i + 1);
            D(i + 1) = cosr * D(i + 1);
            slartg(f, g, cosl, sinl, r);
            D(i) = r;
            f = cosl * E(i) + sinl * D(i + 1);
            D(i + 1) = cosl * D(i + 1) - sinl * E(i);
            if (i < m - 1) {
              g = sinl * E(i + 1);
              E(i + 1) = cosl * E(i + 1);
            }
            W(i - ll + 1) = cosr;
            W(i - ll + 1 + nm1) = sinr;
            W(i - ll + 1 + nm12) = cosl;
            W(i - ll + 1 + nm13) = sinl;
          }
          E(m - 1) = f;
          if (ncvt > 0) slasr_v(true, true, m - ll + 1, ncvt, &W(1), &W(n), vt.sub(ll, 1));
          if (nru > 0) slasr_v(false, true, nru, m - ll + 1, &W(nm12 + 1), &W(nm13 + 1), u.sub(1, ll));
          if (fabsf(E(m - 1)) <= thresh) E(m - 1) = 0.f;
        } else {                       // shifted QR sweep, bottom to top
          float f = (fabsf(D(m)) - shift) * (fsign(1.f, D(m)) + shift / D(m));
          float g = E(m - 1);
          for (int i = m; i >= ll + 1; i--) {
            float cosr, sinr, cosl, sinl;
            slartg(f, g, cosr, sinr, r);
            if (i < m) E(i) = r;
            f = cosr * D(i) + sinr * E(i - 1);
            E(i - 1) = cosr * E(i - 1) - sinr * D(i);
            g = sinr * D(i - 1);
            D(i - 1) = cosr * D(i - 1);
            slartg(f, g, cosl, sinl, r);
            D(i) = r;
            f = cosl * E(i - 1) + sinl * D(i - 1);
            D(i - 1) = cosl * D(i - 1) - sinl * E(i - 1);
            if (i > ll + 1) {
              g = sinl * E(i - 2);
              E(i - 2) = cosl * E(i - 2);
            }
            W(i - ll) = cosr;
            W(i - ll + nm1) = -sinr;
            W(i - ll + nm12) = cosl;
            W(i - ll + nm13) = -sinl;
          }
          E(ll) = f;
          if (fabsf(E(ll)) <= thresh) E(ll) = 0.f;
          if (ncvt > 0) slasr_v(true, false, m - ll + 1, ncvt, &W(nm12 + 1), &W(nm13 + 1), vt.sub(ll, 1));
          if (nru > 0) slasr_v(false, false, nru, m - ll + 1, &W(1), &W(n), u.sub(1, ll));
        }
      }
    }
  }
  // make the singular values positive, then sort them into decreasing order
  for (int i = 1; i <= n; i++) {
    if (D(i) < 0.f) {
      D(i) = -D(i);
      if (ncvt > 0) sscal(ncvt, -1.f, Vec{&vt(i, 1), vt.ld});
    }
  }
  for (int i = 1; i <= n - 1; i++) {
    int isub = 1;
    float smin = D(1);
    for (int j = 2; j <= n + 1 - i; j++) {
      if (D(j) <= smin) {
        isub = j;
        smin = D(j);
      }
    }
    if (isub != n + 1 - i) {
      D(isub) = D(n + 1 - i);
      D(n + 1 - i) = smin;
      if (ncvt > 0) sswap(ncvt, Vec{&vt(isub, 1), vt.ld}, Vec{&vt(n + 1 - i, 1), vt.ld});
      if (nru > 0) sswap(nru, Vec{&u(1, isub), 1}, Vec{&u(1, n + 1 - i), 1});
    }
  }
  return 0;
}

}  // namespace

// The pseudo-inverse (5 x nobs, basis function fastest) of the nobs x 5
// observation matrix of the points xy (nobs (x, y) pairs): columns x, y,
// x*x, x*y, y*y (LSfitting.F90:81-110, Quadratic).  sgesvd('A', 'A')
// (svd.f:1897, the path for m >= n below its QR threshold of
// int(1.6 * n) = 8 rows): bidiagonalise, form U (m x m) and V^T (n x n),
// iterate; then inv = V S^-1 U^T summed over all m columns of U as the
// reference does (LSfitting.F90:110-127).
extern "C" int fimhip_ls_pseudo_inverse(const float *xy, int nobs, float *inv) {
  constexpr int N = 5;
  if (!xy || !inv || nobs < N || nobs > 6) return FIMHIP_EINVAL;
  const int m = nobs;
  float ab[6 * N], ub[6 * 6], vtb[N * N], s[N], e[N], tauq[N], taup[N], work[4 * N + 6];
  memset(ub, 0, sizeof ub);
  memset(vtb, 0, sizeof vtb);
  memset(work, 0, sizeof work);
  Mat a{ab, m}, u{ub, m}, vt{vtb, N};
  for (int i = 1; i <= m; i++) {
    const float x = xy[2 * (i - 1)], y = xy[2 * (i - 1) + 1];
    a(i, 1) = x;
    a(i, 2) = y;
    a(i, 3) = x * x;
    a(i, 4) = x * y;
    a(i, 5) = y * y;
  }
  // sgesvd would rescale a matrix whose largest entry lies outside
  // [sqrt(sfmin)/eps, eps/sqrt(sfmin)] = [2^-40, 2^40] (eps = 2^-23 here)
  float anrm = 0.f;
  for (int j = 1; j <= N; j++)
    for (int i = 1; i <= m; i++) anrm = fmaxf2(anrm, fabsf(a(i, j)));
  if (!(anrm >= ldexpf(1.f, -40) && anrm <= ldexpf(1.f, 40))) return FIMHIP_EINVAL;
  sgebd2(m, N, a, s, e, tauq, taup, work);
  for (int j = 1; j <= N; j++)                  // slacpy('L'): the column reflectors
    for (int i = j; i <= m; i++) u(i, j) = a(i, j);
  sorg2r(m, m, N, u, tauq, work);               // sorgbr('Q', m, m, n) = sorgqr = sorg2r
  for (int j = 1; j <= N; j++)                  // slacpy('U'): the row reflectors
    for (int i = 1; i <= j; i++) vt(i, j) = a(i, j);
  sorgbr_p(N, vt, taup, work);
  if (sbdsqr_u(N, N, m, s, e, vt, u, work) != 0) return FIMHIP_EINVAL;
  // vsigma(i, k) = vt(k, i) / s(k), k <= n, 0 beyond; inv(i, j) = sum_k vsigma(i, k) u(j, k)
  for (int i = 1; i <= N; i++)
    for (int j = 1; j <= m; j++) {
      float acc = 0.f;
      for (int k = 1; k <= m; k++) {
        const float vs = k <= N ? vt(k, i) / s[k - 1] : 0.f;
        acc = acc + vs * u(j, k);
      }
      inv[(i - 1) + (j - 1) * N] = acc;
    }
  return FIMHIP_OK;
}
