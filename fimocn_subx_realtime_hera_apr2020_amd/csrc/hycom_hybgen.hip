// HYCOM's hybrid grid generator (FIMsrc/fim/horizontal/hycom_hybgn1.F90:
// hybgn1, hybgn1_1d, ppmad3, pcwise, equilb; hycom_sigetc.F90: sigocn,
// dsigdt, dsigds) on the GPU in FP64, with the reference's operations in its
// order.  The routine is column-local and its column chains are serial (each
// interface move changes what the next one sees), so one thread takes one
// ocean column (a compacted list of the ocean cells: land leaves no idle
// lanes); its working columns live in private memory.  The velocities and
// the passive tracers are remapped in place: pcwise copies its input column
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
__device__ void y_ppmad3(const double *x0, double *y0, int j, double d) {
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

template <int KM>
struct YCol {
  double pr[KM + 2], tem[KM + 1], sal[KM + 1], dns[KM + 1], pold[KM + 2];
};
template <int KM>
struct YWork {                          // equilb's and pcwise's columns (not live together)
  double w0[KM + 2], w1[KM + 2], w2[KM + 2], w3[KM + 2], w4[KM + 2], w5[KM + 2], w6[KM + 2];
};

// equilb: heat diffusion to inflate thin layers (3 passes), then scenario 4
template <int KM>
__device__ void y_equilb(int kdm, YCol<KM> &c, YWork<KM> &w) {
  const double slak = .25, dffudt = .1;
  double *pnew = w.w0, *pwrk = w.w1, *heatfx = w.w2, *tdz = w.w3, *sdz = w.w4, *twrk = w.w5, *swrk = w.w6;
  const double *thet = c.dns;
  double *pres = c.pr, *temp = c.tem, *saln = c.sal;
  for (int k = 1; k <= kdm + 1; k++) pnew[k] = pwrk[k] = pres[k];
  for (int k = 1; k <= kdm; k++) {
    twrk[k] = temp[k];
    swrk[k] = saln[k];
    tdz[k] = hmax(H_EPSIL, pres[k + 1] - pres[k]) * temp[k];
    sdz[k] = hmax(H_EPSIL, pres[k + 1] - pres[k]) * saln[k];
  }
  heatfx[kdm] = 0.;
  heatfx[1] = 0.;
  for (int iter = 1; iter <= 3; iter++) {
    for (int k = 2; k <= kdm - 1; k++) {
      double hf = 0.;
      if (pwrk[k] > pwrk[1] + H_ONEM && pwrk[k + 1] < pwrk[kdm + 1] - H_ONEM) {
        const double cnv = H_ONEM;
        hf = dffudt * cnv * cnv * hmin(thet[k + 1] - thet[k], thet[k] - thet[k - 1]) /
             hmax(.01 * cnv, pwrk[k + 1] - pwrk[k]);
        const double plimit1 = hmax(pwrk[k] - Y_TENM, .5 * (pwrk[k] + pwrk[k - 1]));
        const double plimit2 = hmin(pwrk[k + 1] + Y_TENM, .5 * (pwrk[k + 1] + pwrk[k + 2]));
        hf = hmin(hmin(hmax(0., hf), hmax(0., (plimit2 - pwrk[k + 1]) * (thet[k + 1] - thet[k]))),
                  hmax(0., (pwrk[k] - plimit1) * (thet[k] - thet[k - 1])));
      }
      heatfx[k] = hf;
    }
    for (int k = 1; k <= kdm - 1; k++) {
      if (thet[k + 1] > thet[k]) {
        pnew[k + 1] = pwrk[k + 1] + (heatfx[k] - heatfx[k + 1]) / (thet[k + 1] - thet[k]);
        const double delp = pnew[k + 1] - pwrk[k + 1];
        const double dtdz = (delp > 0. ? twrk[k + 1] : twrk[k]) * delp;
        const double dsdz = (delp > 0. ? swrk[k + 1] : swrk[k]) * delp;
        tdz[k] = tdz[k] + dtdz;
        tdz[k + 1] = tdz[k + 1] - dtdz;
        sdz[k] = sdz[k] + dsdz;
        sdz[k + 1] = sdz[k + 1] - dsdz;
      }
    }
    for (int k = 1; k <= kdm; k++) {
      if (pnew[k + 1] > pnew[k]) {
        twrk[k] = tdz[k] / (pnew[k + 1] - pnew[k]);
        swrk[k] = sdz[k] / (pnew[k + 1] - pnew[k]);
      }
    }
    for (int k = 1; k <= kdm + 1; k++) pwrk[k] = pnew[k];
  }
  for (int k = 4; k <= kdm - 2; k++) {
    if (pnew[k - 2] < pnew[1] + Y_TENCM || pnew[k + 2] > pnew[kdm + 1] - Y_TENCM) continue;
    const double dp1 = pnew[k - 1] - pnew[k - 2], dp2 = pnew[k] - pnew[k - 1];
    const double dp3 = pnew[k + 1] - pnew[k], dp4 = pnew[k + 2] - pnew[k + 1];
    const double th1 = thet[k - 2], th2 = thet[k - 1], th3 = thet[k], th4 = thet[k + 1];
    if (dp1 > dp2 + dp3 && dp4 > dp2 + dp3) {
      const double goal = .5 * (dp2 + dp3 + hmin(dp1, dp4));
      double dis2 = (goal - dp2 - dp3) * slak, dis4 = dis2;
      if (th2 > th1 && th3 > th2 && th4 > th3) {
        const double ratio = (th4 - th3) / (th2 - th1);
        if (ratio > 1.) dis4 = dis2 / ratio;
        else dis2 = dis4 * ratio;
      }
      pnew[k - 1] = pnew[k - 1] - dis2;
      pnew[k + 1] = pnew[k + 1] + dis4;
      tdz[k - 2] = tdz[k - 2] - dis2 * twrk[k - 2];
      tdz[k - 1] = tdz[k - 1] + dis2 * twrk[k - 2];
      tdz[k] = tdz[k] + dis4 * twrk[k + 1];
      tdz[k + 1] = tdz[k + 1] - dis4 * twrk[k + 1];
      sdz[k - 2] = sdz[k - 2] - dis2 * swrk[k - 2];
      sdz[k - 1] = sdz[k - 1] + dis2 * swrk[k - 2];
      sdz[k] = sdz[k] + dis4 * swrk[k + 1];
      sdz[k + 1] = sdz[k + 1] - dis4 * swrk[k + 1];
    }
  }
  for (int k = 1; k <= kdm; k++) {
    if (pnew[k + 1] - pnew[k] > Y_ONEMU) {
      temp[k] = tdz[k] / (pnew[k + 1] - pnew[k]);
      saln[k] = sdz[k] / (pnew[k + 1] - pnew[k]);
    }
    pres[k] = pnew[k];
  }
}

// pcwise(optn, kdm, pold, fld, pr, fld) without a migration limit: the
// column f (stride st in global memory) remapped from pold to pr in place
template <int KM>
__device__ void y_pcwise(int optn, int kdm, const YCol<KM> &c, YWork<KM> &w, double *f, bool flip) {
  const double athird = 1. / 3., small = 1.e-11;
  double *fco = w.w0, *fcn = w.w1, *delz = w.w2, *fn = w.w3, *a = w.w4, *b = w.w5, *cc = w.w6;
  for (int k = 1; k <= kdm + 1; k++) {
    const double zo = flip ? -c.pold[k] : c.pold[k], zn = flip ? -c.pr[k] : c.pr[k];
    delz[k] = zn - zo;
  }
  for (int k = 1; k <= kdm; k++) {
    const double zo0 = flip ? -c.pold[k] : c.pold[k], zo1 = flip ? -c.pold[k + 1] : c.pold[k + 1];
    const double zn0 = flip ? -c.pr[k] : c.pr[k], zn1 = flip ? -c.pr[k + 1] : c.pr[k + 1];
    fco[k] = hmax(0., zo1 - zo0);
    fcn[k] = hmax(0., zn1 - zn0);
    fn[k] = f[k - 1];
  }
  for (int k = kdm - 1; k >= 1; k--) fn[k] = (fn[k] * fco[k] + fn[k + 1] * small) / (fco[k] + small);
  for (int k = 2; k <= kdm; k++) fn[k] = (fn[k] * fco[k] + fn[k - 1] * small) / (fco[k] + small);
  a[1] = fn[1];
  b[1] = cc[1] = 0.;
  a[kdm] = fn[kdm];
  b[kdm] = cc[kdm] = 0.;
  for (int k = 2; k <= kdm - 1; k++) {
    if (optn == 1) {
      a[k] = fn[k];
      b[k] = cc[k] = 0.;
    } else {
      const double fl = hmax(small, fco[k - 1]), fm = hmax(small, fco[k]), fr = hmax(small, fco[k + 1]);
      const double yl = (fl * fn[k] + fm * fn[k - 1]) / (fl + fm);
      const double yr = (fr * fn[k] + fm * fn[k + 1]) / (fr + fm);
      y_parab(fn[k], yl, yr, a[k], b[k], cc[k]);
    }
  }
  // the flux through interface k from the slab it sweeps; vertdv on the fly
  double vlo = 0.;                      // vertfx(1) = 0
  for (int k = 1; k <= kdm; k++) {
    double vhi = 0.;                    // vertfx(k+1); vertfx(kdm+1) = 0
    const int ki = k + 1;
    if (ki <= kdm) {
      double slab = 0., amount = 0.;
      const double dz = delz[ki];
      if (dz > 0.) {
        for (int lyr = ki;; lyr++) {
          if (slab >= dz) break;
          if (fco[lyr] > 0.) {
            const double dslab = hmin(slab + fco[lyr], dz) - hmin(slab, dz);
            const double dx = dslab / fco[lyr];
            const double fcdx = a[lyr] + b[lyr] * .5 * (dx - 1.) + cc[lyr] * (.25 - dx * (.5 - dx * athird));
            amount = amount + fcdx * dslab;
            slab = slab + dslab;
          }
          if (!(lyr < kdm)) break;
        }
      } else if (dz < 0.) {
        for (int lyr = ki - 1;; lyr--) {  // down to layer 2 only (layer 1 from interface 2)
          if (slab >= -dz) break;
          if (fco[lyr] > 0.) {
            const double dslab = hmin(slab + fco[lyr], -dz) - hmin(slab, -dz);
            const double dx = dslab / fco[lyr];
            const double fcdx = a[lyr] + b[lyr] * .5 * (1. - dx) + cc[lyr] * (.25 - dx * (.5 - dx * athird));
            amount = amount + fcdx * dslab;
            slab = slab + dslab;
          }
          if (!(lyr > 2)) break;
        }
      }
      if (slab != 0.) vhi = -dz * amount / slab;
    }
    const double vertdv = vhi - vlo;
    const double amount = fn[k] * fco[k] - vertdv;
    f[k - 1] = (fn[k] * small + amount) / (small + fcn[k]);
    vlo = vhi;
  }
}

template <int KM>
__global__ void __launch_bounds__(64) hyc_hybgn1(HycHyb h) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= h.ncol) return;
  const int i = h.cols[j], kdm = h.kdm;
  YCol<KM> c;
  YWork<KM> w;
  const double tfreez = -2.2, sigjmp = .01, scalt = -30., scals = 10., slak = 1. / 86400.;
  const double dlt = h.batrop * 30;                 // batrop*bclin_frq
  const double lati = h.latij[i];
  const size_t base = (size_t)i * kdm;
  c.pr[1] = h.pres[(size_t)i * (kdm + 1)];
  for (int k = 1; k <= kdm; k++) {
    c.dns[k] = h.dens[base + k - 1];
    c.tem[k] = h.temp[base + k - 1];
    c.sal[k] = h.saln[base + k - 1];
    c.pr[k + 1] = c.pr[k] + h.dp[base + k - 1];
  }
  double *pr = c.pr, *tem = c.tem, *sal = c.sal, *dns = c.dns;
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
  for (int k = 1; k <= kdm + 1; k++) c.pold[k] = pr[k];

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

  y_equilb<KM>(kdm, c, w);
  for (int k = 1; k <= kdm - 2; k++)
    if (pr[k + 1] < pr[k] || pr[k + 1] > pr[k + 2]) pr[k + 1] = hmax(pr[k], hmin(pr[k + 1], c.pold[k + 2]));

  // the velocities (piecewise constant) and the tracers (parabolic) remapped
  const bool flip = !(c.pold[1] < c.pold[kdm + 1]);
  y_pcwise<KM>(1, kdm, c, w, h.u + base, flip);
  y_pcwise<KM>(1, kdm, c, w, h.v + base, flip);
  for (int nt = 0; nt < h.numtr; nt++) y_pcwise<KM>(3, kdm, c, w, &trc(nt, 1), flip);

  for (int k = 1; k <= kdm; k++) {
    h.dens[base + k - 1] = dns[k];
    h.temp[base + k - 1] = tem[k];
    h.saln[base + k - 1] = sal[k];
    h.dp[base + k - 1] = pr[k + 1] - pr[k];
    h.pres[(size_t)i * (kdm + 1) + k] = pr[k + 1];
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
  g_ktimer = o->kt;
  const dim3 grid((h.ncol + 63) / 64), blk(64);
  if (o->kdm <= 32) KT_LAUNCH("hyc_hybgn1", o->stream, hyc_hybgn1<32><<<grid, blk, 0, o->stream>>>(h));
  else KT_LAUNCH("hyc_hybgn1", o->stream, hyc_hybgn1<64><<<grid, blk, 0, o->stream>>>(h));
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
        (r = o->arena.alloc(&y.dpmin, (size_t)kdm)) || (r = o->arena.alloc(&y.latij, (size_t)nip)))
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
