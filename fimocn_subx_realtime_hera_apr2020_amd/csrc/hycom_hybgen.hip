// HYCOM's hybrid grid generator (FIMsrc/fim/horizontal/hycom_hybgn1.F90:
// hybgn1, hybgn1_1d, ppmad3, pcwise, equilb; hycom_sigetc.F90: sigocn,
// dsigdt, dsigds) on the GPU in FP64, with the reference's operations in its
// order.  The routine is column-local; two kernels split it by the shape of
// its work:
//   hyc_hyb_move   one thread per ocean column (a compacted list of the ocean
//                  cells: land leaves no idle lanes): the sea-floor
//                  absorption, the split of a too-light bottom layer and the
//                  interface moves (hycom_hybgn1.F90:240-758), whose chains
//                  are serial down the column (each move changes what the
//                  next one tests)
//   hyc_hyb_remap  the column's lanes over its layers, its work columns in
//                  LDS: equilb's heat diffusion (per layer and per interface)
//                  and scenario 4, the monotonicity guard, and pcwise's
//                  remapping of the velocities and tracers (fits, sweeps and
//                  updates per layer or interface; its two filling
//                  recurrences and equilb's scenario 4 run on one lane)
// The passive tracers are moved in place: the sea-floor steps change them
// where the reference changes its copy (trold), and pcwise reads its column
// before it writes the result over it.
#include <vector>

#include "hycom_internal.h"

namespace fimhip {
namespace hyc {
namespace {

// hycom_constants.F90:35-41, 54-56, 77-78
constexpr double Y_TENM = 10. * H_ONEM, Y_TENCM = 0.1 * H_ONEM, Y_ONECM = 0.01 * H_ONEM;
constexpr double Y_ONEMM = 0.001 * H_ONEM, Y_ONEMU = 1.e-6 * H_ONEM, Y_RONEM = 1. / H_ONEM;
constexpr double Y_HUUGE = 1.e33, Y_PREF = 1.e7;
constexpr double Y_C2 = 1.470290E-02, Y_C3 = 7.904674E-01, Y_C4 = -6.861022E-03, Y_C5 = -3.051459E-03;
constexpr double Y_C6 = 3.099320E-05, Y_C7 = 3.050299E-05, Y_C8 = 1.046036E-04, Y_C9 = 5.287567E-06;

struct HycHyb {
  const int *cols;                      // the ocean cells, in cell order
  int ncol, nip, kdm, numtr;
  const double *targt, *salmin, *dpmin, *latij;
  double batrop;
  double *dens, *dp, *pres, *temp, *saln, *u, *v, *passv;   // dp, u, v: the level hybgn1 works on
  double *pold;                         // (kdm+1,nip): the interfaces before the moves
};

// hycom_sigetc.F90 sigocn: sigma_1 of t, s (a rational fit at 1000 dbar)
__device__ double y_sigocn(double t, double s) {
  const double p = Y_PREF * 1.e-4;
  const double pn = 9.9984085444849347e2 + t * (7.3471625860981584 + t * (-5.3211231792841769e-2 + 3.6492439109814549e-4 * t)) +
                    s * (2.588057102399139 + -6.7168282786692355e-3 * t + 1.9203202055760151e-3 * s) +
                    p * (1.1798263740430364e-2 + 9.8920219266399117e-8 * t * t + 4.699664277175473e-6 * s +
                         p * (-2.5862187075154352e-8 + -3.2921414007960662e-12 * t * t));
  const double pd = 1.0 + t * (7.2815210113327091e-3 + -4.4787265461983921e-5 * t +
                               t * t * (3.3851002965802430e-7 + 1.3651202389758572e-10 * t)) +
                    s * (1.7632126669040377e-3 + -8.8066583251206474e-6 * t + -1.8832689434804897e-10 * t * t * t +
                         sqrt(hmax(0., s)) * (5.7463776745432097e-6 + 1.4716275472242334e-9 * t * t)) +
                    p * (6.7103246285651894e-6 + p * t * (-2.4461698007024582e-17 * t * t + -9.1534417604289062e-18 * p));
  return pn / pd - 1000.;
}

// the limited parabola of a cell (ppmad3, pcwise, fct3d share it)
__device__ __forceinline__ void y_parab(double y, double yl, double yr, double &a, double &b, double &c) {
  a = 1.5 * y - .25 * (yl + yr);
  b = yr - yl;
  c = 6. * (.5 * (yl + yr) - y);
  if (fabs(yr - yl) < 6. * fabs(.5 * (yl + yr) - y)) {
    if (fabs(yr - yl) > 2. * fabs(.5 * (yl + yr) - y)) {
      if ((yr - yl) * (.5 * (yl + yr) - y) > 0.) {
        a = .25 * (3. * y + yl);
        c = 3. * (y - yl);
        b = c;
      } else {
        a = .25 * (3. * y + yr);
        c = 3. * (y - yr);
        b = -c;
      }
    } else {
      a = y;
      b = 0.;
      c = 0.;
    }
  }
}

// ppmad3 on three cells x0[1..4], values y0[1..3]; one interface moves:
// interface 2 (j == 2) or 3 (j == 3) by d, the others stay
template <typename A>
__device__ void y_ppmad3(const A x0, const A y0, int j, double d) {
  const double onemu = .098, athird = 1. / 3.;
  const double x1 = x0[1], x2 = x0[2], x3 = x0[3], x4 = x0[4];
  const double y1 = y0[1], y2 = y0[2], y3 = y0[3];
  const double dx2 = j == 2 ? d : 0., dx3 = j == 3 ? d : 0.;
  double t1 = y1 * (x2 - x1), t2 = y2 * (x3 - x2), t3 = y3 * (x4 - x3);
  const double dxlft = hmax(onemu, x2 - x1), dxmid = hmax(onemu, x3 - x2), dxrgt = hmax(onemu, x4 - x3);
  const double yl = (dxlft * y2 + dxmid * y1) / (dxlft + dxmid);
  const double yr = (dxrgt * y2 + dxmid * y3) / (dxrgt + dxmid);
  double a, b, c;
  y_parab(y2, yl, yr, a, b, c);
  if (dx2 > 0.) {
    const double wdth = dx2 / (x3 - x2);
    const double slab = (x3 - x2) * wdth * (a + b * .5 * (wdth - 1.) + c * (.25 - wdth * (.5 - wdth * athird)));
    t2 = t2 - slab;
    t1 = t1 + slab;
  }
  if (dx3 < 0.) {
    const double wdth = -dx3 / (x3 - x2);
    const double slab = (x3 - x2) * wdth * (a + b * .5 * (1. - wdth) + c * (.25 - wdth * (.5 - wdth * athird)));
    t2 = t2 - slab;
    t3 = t3 + slab;
  }
  // dx(1) = dx(4) = 0
  double dxnew = x2 - x1 + dx2 - 0.;
  if (dxnew > 0.) y0[1] = t1 / dxnew;
  dxnew = x3 - x2 + dx3 - dx2;
  if (dxnew > 0.) y0[2] = t2 / dxnew;
  dxnew = x4 - x3 + 0. - dx3;
  if (dxnew > 0.) y0[3] = t3 / dxnew;
}

// a column of the thread-per-column kernel in LDS, the block's T threads
// interleaved (adjacent threads, adjacent banks)
template <int T>
struct YS {
  double *p;
  __device__ __forceinline__ double &operator[](int k) const { return p[k * T]; }
  __device__ __forceinline__ YS operator+(int d) const { return YS{p + d * T}; }
  __device__ __forceinline__ YS operator-(int d) const { return YS{p - d * T}; }
};
template <int KM, int T>
__global__ void __launch_bounds__(T) hyc_hyb_move(HycHyb h) {
  __shared__ double sbuf[(KM + 2) + 3 * (KM + 1)][T];
  const int t = threadIdx.x, j = blockIdx.x * T + t;
  if (j >= h.ncol) return;
  const int i = h.cols[j], kdm = h.kdm;
  const YS<T> pr{&sbuf[0][t]}, tem{&sbuf[KM + 2][t]}, sal{&sbuf[2 * KM + 3][t]}, dns{&sbuf[3 * KM + 4][t]};
  const double tfreez = -2.2, sigjmp = .01, scalt = -30., scals = 10., slak = 1. / 86400.;
  const double dlt = h.batrop * 30;                 // batrop*bclin_frq
  const double lati = h.latij[i];
  const size_t base = (size_t)i * kdm;
  pr[1] = h.pres[(size_t)i * (kdm + 1)];
  for (int k = 1; k <= kdm; k++) {
    dns[k] = h.dens[base + k - 1];
    tem[k] = h.temp[base + k - 1];
    sal[k] = h.saln[base + k - 1];
    pr[k + 1] = pr[k] + h.dp[base + k - 1];
  }
  const double *targt = h.targt - 1, *salmin = h.salmin - 1, *dpmin = h.dpmin - 1;   // 1-based
  double q, p_hat = 0.;
  auto slakf = [&]() { return (1. + 9. * hmax(0., 1. - .01 * Y_RONEM * p_hat)) * slak * dlt; };
  auto trc = [&](int nt, int k) -> double & { return h.passv[((size_t)nt * h.nip + i) * kdm + k - 1]; };

  // absorb near-massless layers on the sea floor in the layer above
  int kp = 1;
  for (int k = 2; k <= kdm; k++) {
    if (pr[k] < pr[kdm + 1] - Y_TENCM) {
      kp = k;
    } else {
      const double q1 = hmax(H_EPSIL, pr[k] - pr[kp]), q2 = hmax(0., pr[k + 1] - pr[k]);
      q = q1 / (q1 + q2);
      tem[kp] = tem[kp] * q + tem[k] * (1. - q);
      sal[kp] = sal[kp] * q + sal[k] * (1. - q);
      tem[k] = tem[kp];
      sal[k] = hmax(sal[kp], salmin[k]);
      for (int nt = 0; nt < h.numtr; nt++) {
        trc(nt, kp) = trc(nt, kp) * q + trc(nt, k) * (1. - q);
        trc(nt, k) = trc(nt, kp);
      }
    }
  }
  dns[kp] = y_sigocn(tem[kp], sal[kp]);
  for (int k = kp + 1; k <= kdm; k++) {
    dns[k] = dns[kp];
    pr[k] = pr[kdm + 1];
  }
  // a bottom layer too light: its lower part at the target density, the
  // upper part joined to the layer above
  if (kp != 1) {
    const int k = kp;
    if (!(dns[k] <= hmax(dns[k - 1], targt[k - 1]) || dns[k] >= targt[k])) {
      const double t0 = tem[k], s0 = sal[k];
      const double dsgdt = (Y_C2 + s0 * (Y_C5 + Y_C9 * s0) + 2. * t0 * (Y_C4 + Y_C7 * s0 + 1.5 * Y_C6 * t0)) * scalt;
      const double dsgds = (Y_C3 + 2. * s0 * (Y_C8 + Y_C9 * t0) + t0 * (Y_C5 + t0 * Y_C7)) * scals;
      q = 1. / (dsgdt + dsgds);
      const double sal_lo = hmin(s0 + (targt[k] - dns[k]) * q * scals, hmax(sal[k - 1], s0 + .01 * scals));
      const double tem_lo = hmax(t0 + (targt[k] - dns[k]) * q * scalt, tfreez);
      const double rho_lo = y_sigocn(tem_lo, sal_lo);
      const double rho_up = hmax(dns[k - 1], hmin(targt[k - 1], dns[k]));
      q = (rho_lo - dns[k]) / hmax(H_EPSIL, rho_lo - rho_up);
      if (!(q < 0. || q > 1.)) {
        p_hat = pr[k] * (1. - q) + pr[k + 1] * q;
        q = (pr[k + 1] - p_hat) / hmax(H_EPSIL, p_hat - pr[k]);
        const double sal_up = s0 * (1. + q) - sal_lo * q;
        const double tem_up = t0 * (1. + q) - tem_lo * q;
        q = (p_hat - pr[k]) / hmax(p_hat - pr[k - 1], H_EPSIL);
        sal[k - 1] = sal_up * q + sal[k - 1] * (1. - q);
        tem[k - 1] = tem_up * q + tem[k - 1] * (1. - q);
        for (int nt = 0; nt < h.numtr; nt++) trc(nt, k - 1) = trc(nt, k) * q + trc(nt, k - 1) * (1. - q);
        dns[k - 1] = y_sigocn(tem[k - 1], sal[k - 1]);
        tem[k] = tem_lo;
        sal[k] = sal_lo;
        dns[k] = rho_lo;
        pr[k] = p_hat;
      }
    }
  }
  const size_t pb = (size_t)i * (kdm + 1);
  for (int k = 1; k <= kdm + 1; k++) h.pold[pb + k - 1] = pr[k];     // pcwise's old interfaces

  // restore isopycnic conditions by moving the interfaces
  double dpsum = 0., dp0 = Y_HUUGE;
  for (int k = 1; k <= kdm; k++) {
    const double dp0abv = dp0;
    dp0 = dpmin[k] * H_ONEM;
    if (k > 1) dp0 = dp0 * hmax(.6, hmin(1., (fabs(lati) + 5.) * .04));
    if (4 * k < kdm) dp0 = dp0 * hmin(1., pr[kdm + 1] / (200. * H_ONEM) + .4);
    dpsum = dpsum + dp0;
    bool lower = false;                 // the move of the lower interface (the reference's label 5)
    if (k == 1) {
      p_hat = dp0;
      if (p_hat > pr[2]) {              // layer 1 too thin
        p_hat = hmin(p_hat, pr[2] + hmax(Y_ONECM, slakf() * (p_hat - pr[2])));
        lower = true;
      } else if (p_hat < pr[2]) {       // too thick: its water into layer 2
        p_hat = hmax(p_hat, pr[2] + hmin(-Y_ONECM, slakf() * (p_hat - pr[2])));
        q = (pr[2] - p_hat) / hmax(pr[3] - p_hat, H_EPSIL);
        sal[2] = sal[2] * (1. - q) + sal[1] * q;
        tem[2] = tem[2] * (1. - q) + tem[1] * q;
        dns[2] = y_sigocn(tem[2], sal[2]);
        pr[2] = p_hat;
      }
    } else {
      if (pr[k] == pr[kdm + 1]) dns[k] = hmax(targt[k], dns[k]);
      if (pr[k] > pr[kdm + 1] - Y_ONECM) continue;
      p_hat = dpsum;
      if (k < kdm && p_hat > pr[k + 1]) {           // lower interface too close to the surface
        p_hat = hmin(p_hat, pr[k + 1] + hmax(Y_ONECM, slakf() * (p_hat - pr[k + 1])));
        lower = true;
      } else if (fabs(dns[k] - targt[k]) < .1 * sigjmp) {
        continue;
      } else if (dns[k] <= targt[k]) {              // too light: take water from below
        if (k >= kdm || pr[k + 1] > pr[kdm + 1] - Y_ONEMM) continue;
        q = (dns[k] - targt[k]) / hmax(dns[k + 1] - targt[k], sigjmp * 10.);
        p_hat = pr[k + 1] * (1. - q) + pr[k] * q;
        p_hat = hmin(p_hat, hmax(.5 * (pr[k] + pr[k + 2]), 2. * pr[k + 1] - pr[k]));
        p_hat = hmin(p_hat, pr[k + 1] + hmax(Y_ONECM, slakf() * (p_hat - pr[k + 1])));
        lower = true;
      } else {                                      // too dense: take water from above
        if (k != 2) {
          q = (targt[k] - dns[k]) / hmax(targt[k] - dns[k - 1], sigjmp * 10.);
          p_hat = pr[k] * (1. - q) + pr[k + 1] * q;
          p_hat = hmax(p_hat, pr[k - 1] + dp0abv);
          p_hat = hmin(p_hat, .5 * (pr[k - 1] + pr[k + 1]));
          if (p_hat < pr[k]) {                      // upper interface up
            p_hat = hmax(hmax(p_hat, pr[k - 1]), pr[k] + hmin(-Y_ONECM, slakf() * (p_hat - pr[k])));
            if (fabs(dns[k - 1] - targt[k - 1]) > .1 * sigjmp) {
              y_ppmad3(pr + k - 3, sal + k - 3, 3, p_hat - pr[k]);
              y_ppmad3(pr + k - 3, tem + k - 3, 3, p_hat - pr[k]);
              dns[k - 1] = y_sigocn(tem[k - 1], sal[k - 1]);
              dns[k] = y_sigocn(tem[k], sal[k]);
            } else {
              q = (pr[k] - p_hat) / hmax(pr[k + 1] - p_hat, H_EPSIL);
              sal[k] = sal[k] * (1. - q) + sal[k - 1] * q;
              tem[k] = tem[k] * (1. - q) + tem[k - 1] * q;
              dns[k] = y_sigocn(tem[k], sal[k]);
            }
            pr[k] = p_hat;
          } else if (p_hat > pr[k]) {               // upper interface down
            p_hat = hmin(hmin(p_hat, pr[k + 1]), pr[k] + hmax(Y_ONECM, slakf() * (p_hat - pr[k])));
            if (k < kdm) {
              y_ppmad3(pr + k - 2, sal + k - 2, 2, p_hat - pr[k]);
              y_ppmad3(pr + k - 2, tem + k - 2, 2, p_hat - pr[k]);
              dns[k - 1] = y_sigocn(tem[k - 1], sal[k - 1]);
              dns[k] = y_sigocn(tem[k], sal[k]);
            } else {
              q = (p_hat - pr[k]) / hmax(p_hat - pr[k - 1], H_EPSIL);
              sal[k - 1] = sal[k - 1] * (1. - q) + sal[k] * q;
              tem[k - 1] = tem[k - 1] * (1. - q) + tem[k] * q;
              dns[k - 1] = y_sigocn(tem[k - 1], sal[k - 1]);
            }
            pr[k] = p_hat;
          }
        }
        p_hat = pr[k] + dpmin[2] * H_ONEM;          // inflate layer k downward?
        if (k < kdm && pr[k + 1] < pr[kdm + 1] - Y_ONEMM && pr[k + 1] < p_hat) {
          p_hat = hmin(p_hat, pr[k + 1] + hmax(Y_ONECM, slakf() * (p_hat - pr[k + 1])));
          lower = true;
        }
      }
    }
    if (!lower) continue;
    p_hat = hmin(p_hat, pr[k + 2]);
    if (p_hat > pr[k + 1] + Y_ONEMM) {
      if (k < kdm - 1 && fabs(dns[k + 1] - targt[k + 1]) > .1 * sigjmp) {
        y_ppmad3(pr + k - 1, sal + k - 1, 2, p_hat - pr[k + 1]);
        y_ppmad3(pr + k - 1, tem + k - 1, 2, p_hat - pr[k + 1]);
        dns[k] = y_sigocn(tem[k], sal[k]);
        dns[k + 1] = y_sigocn(tem[k + 1], sal[k + 1]);
      } else {
        q = (p_hat - pr[k + 1]) / hmax(p_hat - pr[k], H_EPSIL);
        sal[k] = sal[k] * (1. - q) + sal[k + 1] * q;
        tem[k] = tem[k] * (1. - q) + tem[k + 1] * q;
        dns[k] = y_sigocn(tem[k], sal[k]);
      }
      pr[k + 1] = p_hat;
    }
  }

  // handed to hyc_hyb_remap: the moved interfaces and the layers
  for (int k = 1; k <= kdm + 1; k++) h.pres[pb + k - 1] = pr[k];
  for (int k = 1; k <= kdm; k++) {
    h.dens[base + k - 1] = dns[k];
    h.temp[base + k - 1] = tem[k];
    h.saln[base + k - 1] = sal[k];
  }
}

// equilb, the monotonicity guard and the pcwise remapping of one column on
// LPC lanes: lane l holds layer k = l + 1 and the interface below it (k + 1)
template <int LPC>
__global__ void __launch_bounds__(256) hyc_hyb_remap(HycHyb h) {
  constexpr int CPB = 256 / LPC, N = LPC + 2;
  __shared__ double sP[CPB][N], sW[CPB][N], sO[CPB][N], sT[CPB][N], sTW[CPB][N], sSW[CPB][N], sTD[CPB][N],
      sSD[CPB][N], sH[CPB][N];
  const int cc = threadIdx.x / LPC, l = threadIdx.x % LPC;
  const int j = blockIdx.x * CPB + cc;
  if (j >= h.ncol) return;
  const int i = h.cols[j], kdm = h.kdm, k = l + 1;
  const bool lay = k <= kdm;
  double *P = sP[cc], *W = sW[cc], *PO = sO[cc], *TH = sT[cc], *TW = sTW[cc], *SW = sSW[cc], *TD = sTD[cc];
  double *SD = sSD[cc], *HF = sH[cc];
  const size_t base = (size_t)i * kdm, pb = (size_t)i * (kdm + 1);
  for (int m = k; m <= kdm + 1; m += LPC) {
    P[m] = W[m] = h.pres[pb + m - 1];
    PO[m] = h.pold[pb + m - 1];
  }
  double tem = 0., sal = 0.;
  if (lay) {
    TH[k] = h.dens[base + k - 1];
    tem = h.temp[base + k - 1];
    sal = h.saln[base + k - 1];
    TW[k] = tem;
    SW[k] = sal;
  }
  wave_sync();
  // equilb (hycom_hybgn1.F90:1289-1752): heat diffusion, three passes
  const double slak = .25, dffudt = .1;
  double td = 0., sd = 0.;
  if (lay) {
    td = hmax(H_EPSIL, P[k + 1] - P[k]) * tem;
    sd = hmax(H_EPSIL, P[k + 1] - P[k]) * sal;
  }
  for (int iter = 1; iter <= 3; iter++) {
    double hf = 0.;
    if (k >= 2 && k <= kdm - 1 && W[k] > W[1] + H_ONEM && W[k + 1] < W[kdm + 1] - H_ONEM) {
      const double cnv = H_ONEM;
      hf = dffudt * cnv * cnv * hmin(TH[k + 1] - TH[k], TH[k] - TH[k - 1]) / hmax(.01 * cnv, W[k + 1] - W[k]);
      const double plimit1 = hmax(W[k] - Y_TENM, .5 * (W[k] + W[k - 1]));
      const double plimit2 = hmin(W[k + 1] + Y_TENM, .5 * (W[k + 1] + W[k + 2]));
      hf = hmin(hmin(hmax(0., hf), hmax(0., (plimit2 - W[k + 1]) * (TH[k + 1] - TH[k]))),
                hmax(0., (W[k] - plimit1) * (TH[k] - TH[k - 1])));
    }
    if (lay) HF[k] = hf;
    wave_sync();
    // the interface below layer k moves by the flux difference; the heat
    // and salt it sweeps leave one layer for the other
    double dt = 0., ds = 0.;
    int moved = 0;
    if (k <= kdm - 1 && TH[k + 1] > TH[k]) {
      moved = 1;
      const double pn = W[k + 1] + (HF[k] - HF[k + 1]) / (TH[k + 1] - TH[k]);
      const double delp = pn - W[k + 1];
      dt = (delp > 0. ? TW[k + 1] : TW[k]) * delp;
      ds = (delp > 0. ? SW[k + 1] : SW[k]) * delp;
      P[k + 1] = pn;
    }
    const double dtu = __shfl_up(dt, 1, LPC), dsu = __shfl_up(ds, 1, LPC);
    const int movu = __shfl_up(moved, 1, LPC);
    if (lay) {                          // the interface above layer k is handled first
      if (k > 1 && movu) {
        td = td - dtu;
        sd = sd - dsu;
      }
      if (moved) {
        td = td + dt;
        sd = sd + ds;
      }
    }
    wave_sync();
    if (lay && P[k + 1] > P[k]) {
      TW[k] = td / (P[k + 1] - P[k]);
      SW[k] = sd / (P[k + 1] - P[k]);
    }
    wave_sync();
    for (int m = k; m <= kdm + 1; m += LPC) W[m] = P[m];
    wave_sync();
  }
  if (lay) {
    TD[k] = td;
    SD[k] = sd;
  }
  wave_sync();
  if (l == 0) {                         // scenario 4: thick-(thin-thin)-thick, down the column
    for (int kk = 4; kk <= kdm - 2; kk++) {
      if (P[kk - 2] < P[1] + Y_TENCM || P[kk + 2] > P[kdm + 1] - Y_TENCM) continue;
      const double dp1 = P[kk - 1] - P[kk - 2], dp2 = P[kk] - P[kk - 1];
      const double dp3 = P[kk + 1] - P[kk], dp4 = P[kk + 2] - P[kk + 1];
      const double th1 = TH[kk - 2], th2 = TH[kk - 1], th3 = TH[kk], th4 = TH[kk + 1];
      if (dp1 > dp2 + dp3 && dp4 > dp2 + dp3) {
        const double goal = .5 * (dp2 + dp3 + hmin(dp1, dp4));
        double dis2 = (goal - dp2 - dp3) * slak, dis4 = dis2;
        if (th2 > th1 && th3 > th2 && th4 > th3) {
          const double ratio = (th4 - th3) / (th2 - th1);
          if (ratio > 1.) dis4 = dis2 / ratio;
          else dis2 = dis4 * ratio;
        }
        P[kk - 1] = P[kk - 1] - dis2;
        P[kk + 1] = P[kk + 1] + dis4;
        TD[kk - 2] = TD[kk - 2] - dis2 * TW[kk - 2];
        TD[kk - 1] = TD[kk - 1] + dis2 * TW[kk - 2];
        TD[kk] = TD[kk] + dis4 * TW[kk + 1];
        TD[kk + 1] = TD[kk + 1] - dis4 * TW[kk + 1];
        SD[kk - 2] = SD[kk - 2] - dis2 * SW[kk - 2];
        SD[kk - 1] = SD[kk - 1] + dis2 * SW[kk - 2];
        SD[kk] = SD[kk] + dis4 * SW[kk + 1];
        SD[kk + 1] = SD[kk + 1] - dis4 * SW[kk + 1];
      }
    }
  }
  wave_sync();
  if (lay && P[k + 1] - P[k] > Y_ONEMU) {
    tem = TD[k] / (P[k + 1] - P[k]);
    sal = SD[k] / (P[k + 1] - P[k]);
  }
  wave_sync();
  // the monotonicity guard of hybgn1_1d (hycom_hybgn1.F90:768-776)
  if (l == 0)
    for (int kk = 1; kk <= kdm - 2; kk++)
      if (P[kk + 1] < P[kk] || P[kk + 1] > P[kk + 2]) P[kk + 1] = hmax(P[kk], hmin(P[kk + 1], PO[kk + 2]));
  wave_sync();
  // pcwise (hycom_hybgn1.F90:981-1244) with no migration limit: the
  // velocities piecewise constant, the tracers piecewise parabolic
  const double athird = 1. / 3., small = 1.e-11;
  const bool flip = !(PO[1] < PO[kdm + 1]);
  auto z = [&](double x) { return flip ? -x : x; };
  double fco = 0., fcn = 0., dz = 0.;
  if (lay) {
    fco = hmax(0., z(PO[k + 1]) - z(PO[k]));
    fcn = hmax(0., z(P[k + 1]) - z(P[k]));
  }
  if (k <= kdm - 1) dz = z(P[k + 1]) - z(PO[k + 1]);       // delz of the interface below layer k
  double *FC = TW, *FN = W, *A = SW, *B = TD, *CQ = SD;    // equilb's columns are free now
  if (lay) FC[k] = fco;
  const int nfield = 2 + h.numtr;
  for (int f = 0; f < nfield; f++) {
    double *g = f == 0 ? h.u + base : f == 1 ? h.v + base : h.passv + ((size_t)(f - 2) * h.nip + i) * kdm;
    wave_sync();
    if (lay) FN[k] = g[k - 1];
    wave_sync();
    if (l == 0) {                       // zero-width cells take their neighbours' values
      for (int kk = kdm - 1; kk >= 1; kk--) FN[kk] = (FN[kk] * FC[kk] + FN[kk + 1] * small) / (FC[kk] + small);
      for (int kk = 2; kk <= kdm; kk++) FN[kk] = (FN[kk] * FC[kk] + FN[kk - 1] * small) / (FC[kk] + small);
    }
    wave_sync();
    const double fn = lay ? FN[k] : 0.;
    if (lay) {
      double a = fn, b = 0., c = 0.;
      if (f >= 2 && k > 1 && k < kdm) {
        const double fl = hmax(small, FC[k - 1]), fm = hmax(small, fco), fr = hmax(small, FC[k + 1]);
        const double yl = (fl * fn + fm * FN[k - 1]) / (fl + fm);
        const double yr = (fr * fn + fm * FN[k + 1]) / (fr + fm);
        y_parab(fn, yl, yr, a, b, c);
      }
      A[k] = a;
      B[k] = b;
      CQ[k] = c;
    }
    wave_sync();
    double vhi = 0.;                    // vertfx of the interface below layer k
    if (k <= kdm - 1) {
      const int ki = k + 1;
      double slab = 0., amount = 0.;
      if (dz > 0.) {
        for (int lyr = ki;; lyr++) {
          if (slab >= dz) break;
          const double fc = FC[lyr];
          if (fc > 0.) {
            const double dslab = hmin(slab + fc, dz) - hmin(slab, dz);
            const double dx = dslab / fc;
            const double fcdx = A[lyr] + B[lyr] * .5 * (dx - 1.) + CQ[lyr] * (.25 - dx * (.5 - dx * athird));
            amount = amount + fcdx * dslab;
            slab = slab + dslab;
          }
          if (!(lyr < kdm)) break;
        }
      } else if (dz < 0.) {
        for (int lyr = ki - 1;; lyr--) {  // down to layer 2 (layer 1 from interface 2 only)
          if (slab >= -dz) break;
          const double fc = FC[lyr];
          if (fc > 0.) {
            const double dslab = hmin(slab + fc, -dz) - hmin(slab, -dz);
            const double dx = dslab / fc;
            const double fcdx = A[lyr] + B[lyr] * .5 * (1. - dx) + CQ[lyr] * (.25 - dx * (.5 - dx * athird));
            amount = amount + fcdx * dslab;
            slab = slab + dslab;
          }
          if (!(lyr > 2)) break;
        }
      }
      if (slab != 0.) vhi = -dz * amount / slab;
    }
    const double vup = __shfl_up(vhi, 1, LPC);
    if (lay) {
      const double vertdv = vhi - (k == 1 ? 0. : vup);
      const double amount = fn * fco - vertdv;
      g[k - 1] = (fn * small + amount) / (small + fcn);
    }
  }
  if (lay) {
    h.temp[base + k - 1] = tem;
    h.saln[base + k - 1] = sal;
    h.dp[base + k - 1] = P[k + 1] - P[k];
    h.pres[pb + k] = P[k + 1];
  }
}

}  // namespace
}  // namespace hyc
}  // namespace fimhip

using namespace fimhip;
using namespace fimhip::hyc;

static int hyb_launch(fimhip_ocn *o, int leap, int with_tracers) {
  if (leap < 1 || leap > 2 || !o->hy.cols || !(o->mo.batrop > 0.)) return FIMHIP_EINVAL;
  if (with_tracers && o->tr.numtr <= 0) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  const size_t cz = (size_t)o->nip * o->kdm;
  HycHyb h{};
  h.cols = o->hy.cols;
  h.ncol = o->hy.ncol;
  h.nip = o->nip;
  h.kdm = o->kdm;
  h.numtr = with_tracers ? o->tr.numtr : 0;
  h.targt = o->hy.targt;
  h.salmin = o->hy.salmin;
  h.dpmin = o->hy.dpmin;
  h.latij = o->hy.latij;
  h.batrop = o->mo.batrop;
  h.dens = o->hy.dens;
  h.dp = o->s.dp + (leap - 1) * cz;
  h.pres = o->s.pres;
  h.temp = o->tr.temp;
  h.saln = o->tr.saln;
  h.u = o->ed.uclin + (leap - 1) * cz;
  h.v = o->ed.vclin + (leap - 1) * cz;
  h.passv = o->tr.passv;
  if (h.ncol == 0) return FIMHIP_OK;
  h.pold = o->hy.pold;
  g_ktimer = o->kt;
  const dim3 b2(256);
  if (o->kdm <= 32) {
    KT_LAUNCH("hyc_hyb_move", o->stream, hyc_hyb_move<32, 32><<<dim3((h.ncol + 31) / 32), dim3(32), 0, o->stream>>>(h));
    KT_LAUNCH("hyc_hyb_remap", o->stream, hyc_hyb_remap<32><<<dim3((h.ncol + 7) / 8), b2, 0, o->stream>>>(h));
  } else {
    KT_LAUNCH("hyc_hyb_move", o->stream, hyc_hyb_move<64, 16><<<dim3((h.ncol + 15) / 16), dim3(16), 0, o->stream>>>(h));
    KT_LAUNCH("hyc_hyb_remap", o->stream, hyc_hyb_remap<64><<<dim3((h.ncol + 3) / 4), b2, 0, o->stream>>>(h));
  }
  g_ktimer = nullptr;
  HIPCHK(hipGetLastError());
  return FIMHIP_OK;
}

extern "C" int fimhip_ocn_set_hybgen(fimhip_ocn *o, const double *targt, const double *salmin, const double *dpmin,
                                     const double *latij) {
  if (!o || !targt || !salmin || !dpmin || !latij) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  HycHybState &y = o->hy;
  const int nip = o->nip, kdm = o->kdm;
  int r;
  if (!y.cols) {
    std::vector<int> cols;
    for (int i = 0; i < nip; i++)
      if (o->wet[i] > 0) cols.push_back(i);
    int *dc = nullptr;
    if ((r = o->arena.alloc(&dc, cols.size() + 1))) return r;
    if ((r = o->arena.alloc(&y.targt, (size_t)kdm)) || (r = o->arena.alloc(&y.salmin, (size_t)kdm)) ||
        (r = o->arena.alloc(&y.dpmin, (size_t)kdm)) || (r = o->arena.alloc(&y.latij, (size_t)nip)) ||
        (r = o->arena.alloc(&y.pold, (size_t)nip * (kdm + 1))))
      return r;
    HIPCHK(hipMemcpyAsync(dc, cols.data(), cols.size() * sizeof(int), hipMemcpyHostToDevice, o->stream));
    HIPCHK(hipStreamSynchronize(o->stream));
    y.cols = dc;
    y.ncol = (int)cols.size();
  }
  HIPCHK(hipMemcpyAsync(y.targt, targt, kdm * sizeof(double), hipMemcpyHostToDevice, o->stream));
  HIPCHK(hipMemcpyAsync(y.salmin, salmin, kdm * sizeof(double), hipMemcpyHostToDevice, o->stream));
  HIPCHK(hipMemcpyAsync(y.dpmin, dpmin, kdm * sizeof(double), hipMemcpyHostToDevice, o->stream));
  HIPCHK(hipMemcpyAsync(y.latij, latij, nip * sizeof(double), hipMemcpyHostToDevice, o->stream));
  HIPCHK(hipStreamSynchronize(o->stream));
  return FIMHIP_OK;
}

// hybgn1 on the device-resident arrays, asynchronous
extern "C" int fimhip_ocn_hybgn1(fimhip_ocn *o, int leap, int with_tracers) {
  if (!o) return FIMHIP_EINVAL;
  return hyb_launch(o, leap, with_tracers);
}

extern "C" int fimhip_ocn_hybgn1_r8(fimhip_ocn *o, int nstep, int leap, const double *targt, double *dens, double *dp,
                                    double *pres, double *temp, double *saln, double *uvel, double *vvel,
                                    const double *latij, double *passv_tr) {
  (void)nstep;
  if (!o || !targt || !dens || !dp || !pres || !temp || !saln || !uvel || !vvel || !latij) return FIMHIP_EINVAL;
  if (!o->hy.cols || leap < 1 || leap > 2 || (passv_tr && o->tr.numtr <= 0)) return FIMHIP_EINVAL;
  HIPCHK(hipSetDevice(o->device));
  HIPCHK(hipMemcpyAsync(o->hy.targt, targt, o->kdm * sizeof(double), hipMemcpyHostToDevice, o->stream));
  HIPCHK(hipMemcpyAsync(o->hy.latij, latij, o->nip * sizeof(double), hipMemcpyHostToDevice, o->stream));
  double *io[] = {dens, dp, pres, temp, saln, uvel, vvel, passv_tr};
  const int fi[] = {FIMHIP_OCN_DENS, FIMHIP_OCN_DP, FIMHIP_OCN_PRES, FIMHIP_OCN_TEMP, FIMHIP_OCN_SALN,
                    FIMHIP_OCN_UCLIN, FIMHIP_OCN_VCLIN, FIMHIP_OCN_PASSV};
  const int nio = passv_tr ? 8 : 7;
  for (int j = 0; j < nio; j++) {
    size_t n = 0;
    double *d = ocn_field(o, fi[j], &n);
    if (!d) return FIMHIP_EINVAL;
    HIPCHK(hipMemcpyAsync(d, io[j], n * sizeof(double), hipMemcpyHostToDevice, o->stream));
  }
  const int r = hyb_launch(o, leap, passv_tr ? 1 : 0);
  if (r) return r;
  for (int j = 0; j < nio; j++) {
    size_t n = 0;
    double *d = ocn_field(o, fi[j], &n);
    HIPCHK(hipMemcpyAsync(io[j], d, n * sizeof(double), hipMemcpyDeviceToHost, o->stream));
  }
  HIPCHK(hipStreamSynchronize(o->stream));
  return FIMHIP_OK;
}
