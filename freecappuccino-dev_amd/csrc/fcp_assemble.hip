// fcp_assemble.hip -- matrix/source assembly and cell gradients.
//
// The reference scatters every face into its owner and neighbour rows in
// ascending face order (laplacian.f90:40-102, gradients.f90:1647-1690).  Each
// kernel here runs one thread per cell and walks the cell's faces in
// ascending face id, repeating the per-face arithmetic of the reference, so
// every row is produced without atomics and bit-equal to the reference.
#include "fcp_internal.h"

#include <algorithm>

namespace {

constexpr int BLK = 256;

// Cell of this thread, XCD-chunked: workgroups are dispatched to the eight
// XCDs round-robin (workgroup b on XCD b % 8), each XCD with its own L2.
// XCD x takes the x-th eighth of the cells, so the face and cell data that
// neighbouring cells share -- those one mesh plane away included -- are
// fetched into one L2 rather than into two.  Launch xcd_grid(m) workgroups.
__device__ __forceinline__ int xcd_cell(int m) {
  const int chunk = ((m + 8 * BLK - 1) / (8 * BLK)) * BLK;
  return (int)(blockIdx.x & 7) * chunk + (int)(blockIdx.x >> 3) * BLK + (int)threadIdx.x;
}
int xcd_grid(int m) {
  const int chunk = ((m + 8 * BLK - 1) / (8 * BLK)) * BLK;
  return m > 0 ? 8 * (chunk / BLK) : 1;
}

struct GeomP {
  const double *xc, *yc, *zc, *vol, *arx, *ary, *arz, *xf, *yf, *zf, *facint, *Df;
};

// laplacian.f90:1-151.  mu (per cell) or mu_f (per face) selects the form;
// bc_kind == nullptr treats every boundary face as Dirichlet.
__global__ void k_laplacian(int n, int Fi, const int *__restrict__ cf_ptr,
                            const int *__restrict__ cf, const int *__restrict__ owner,
                            const int *__restrict__ neighbour, GeomP g,
                            const double *__restrict__ mu, const double *__restrict__ mu_f,
                            const unsigned char *__restrict__ bc_kind,
                            const double *__restrict__ phi, const int *__restrict__ slot_pn,
                            const int *__restrict__ slot_np, const int *__restrict__ diag,
                            double *__restrict__ a, double *__restrict__ su) {
  const int c = xcd_cell(n);
  if (c >= n) return;
  double ad = 0.0;  // a(diag) starts from the cleared matrix (:35)
  double s = su[c];
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int f = cf[k];
    if (f < Fi) {
      const int p = owner[f], q = neighbour[f];
      double cap;
      if (mu_f) {
        cap = mu_f[f] * g.Df[f];
      } else {  // faceFluxLaplacian :113-151
        const double fxn = g.facint[f], fxp = 1.0 - g.facint[f];
        const double xpn = g.xc[q] - g.xc[p], ypn = g.yc[q] - g.yc[p], zpn = g.zc[q] - g.zc[p];
        const double ax = g.arx[f], ay = g.ary[f], az = g.arz[f];
        const double smdpn = ((ax * ax + ay * ay) + az * az) / ((ax * xpn + ay * ypn) + az * zpn);
        cap = (fxp * mu[p] + fxn * mu[q]) * smdpn;
      }
      const int slot = (c == p) ? slot_pn[f] : slot_np[f];
      if (slot >= 0) a[slot] = cap;
      ad = ad - cap;
    } else {
      if (bc_kind && bc_kind[f - Fi] != 0) continue;
      const double ax = g.arx[f], ay = g.ary[f], az = g.arz[f];
      const double are = sqrt((ax * ax + ay * ay) + az * az);
      const double nxf = ax / are, nyf = ay / are, nzf = az / are;
      const double dfn = ((g.xf[f] - g.xc[c]) * nxf + (g.yf[f] - g.yc[c]) * nyf) +
                         (g.zf[f] - g.zc[c]) * nzf;
      const double dcoef = (mu_f ? mu_f[f] : mu[c]) * are / dfn;
      ad = ad - dcoef;
      s = s - dcoef * phi[n + (f - Fi)];
    }
  }
  a[diag[c]] = ad;
  su[c] = s;
}

// Explicit non-orthogonal correction of facefluxsc (scalar_fluxes.f90:99-109)
// with zero mass flux: su(P) += s, su(N) -= s (temperature.f90:153-154).
__global__ void k_nonorth(int n, int Fi, const int *__restrict__ cf_ptr, const int *__restrict__ cf,
                          const int *__restrict__ owner, const int *__restrict__ neighbour, GeomP g,
                          const double *__restrict__ gam, const double *__restrict__ gr,
                          double *__restrict__ su) {
  const int c = xcd_cell(n);
  if (c >= n) return;
  double s = su[c];
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int f = cf[k];
    if (f >= Fi) continue;
    const int p = owner[f], q = neighbour[f];
    const double fxn = g.facint[f], fxp = 1.0 - g.facint[f];
    const double xpn = g.xc[q] - g.xc[p], ypn = g.yc[q] - g.yc[p], zpn = g.zc[q] - g.zc[p];
    const double Dff = g.Df[f];
    double dfixi = gr[3 * p] * fxp + gr[3 * q] * fxn;
    double dfiyi = gr[3 * p + 1] * fxp + gr[3 * q + 1] * fxn;
    double dfizi = gr[3 * p + 2] * fxp + gr[3 * q + 2] * fxn;
    dfixi = dfixi * (g.arx[f] - Dff * xpn);
    dfiyi = dfiyi * (g.ary[f] - Dff * ypn);
    dfizi = dfizi * (g.arz[f] - Dff * zpn);
    const double fdfie = gam[f] * ((dfixi + dfiyi) + dfizi);
    s = (c == p) ? s + fdfie : s - fdfie;
  }
  su[c] = s;
}

// grad_gauss (gradients.f90:1607-1715); boundary rows of dudxi are zero.
__global__ void k_grad_gauss(int n, int Fi, int ntot, const int *__restrict__ cf_ptr,
                             const int *__restrict__ cf, const int *__restrict__ owner,
                             const int *__restrict__ neighbour, GeomP g,
                             const double *__restrict__ u, double *__restrict__ dudxi) {
  const int c = xcd_cell(ntot);
  if (c >= ntot) return;
  if (c >= n) {
    dudxi[3 * (size_t)c] = 0.0;
    dudxi[3 * (size_t)c + 1] = 0.0;
    dudxi[3 * (size_t)c + 2] = 0.0;
    return;
  }
  double gx = 0.0, gy = 0.0, gz = 0.0;
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int f = cf[k];
    if (f < Fi) {
      const int p = owner[f], q = neighbour[f];
      const double fie = u[p] + (u[q] - u[p]) * g.facint[f];
      const double dfxe = fie * g.arx[f], dfye = fie * g.ary[f], dfze = fie * g.arz[f];
      if (c == p) {
        gx = gx + dfxe;
        gy = gy + dfye;
        gz = gz + dfze;
      } else {
        gx = gx - dfxe;
        gy = gy - dfye;
        gz = gz - dfze;
      }
    } else {  // gradbc :1696-1715
      const double fi = u[n + (f - Fi)];
      gx = gx + fi * g.arx[f];
      gy = gy + fi * g.ary[f];
      gz = gz + fi * g.arz[f];
    }
  }
  const double volr = 1.0 / g.vol[c];
  dudxi[3 * (size_t)c] = gx * volr;
  dudxi[3 * (size_t)c + 1] = gy * volr;
  dudxi[3 * (size_t)c + 2] = gz * volr;
}

// One sweep of fvxGradient.f90's Gauss gradient: the inner-face value is
// corrected by the old gradient dfo (gradco :1761-1817; dfo == nullptr is
// the zero gradient of grad_gauss's first sweep, :1589-1591), boundary faces
// by gradbc (:1819-1838).  Each cell gathers its faces in ascending order,
// which is the order of the reference's face loops for that cell.
__global__ void k_grad_fvx(int n, int Fi, int ntot, const int *__restrict__ cf_ptr,
                           const int *__restrict__ cf, const int *__restrict__ owner,
                           const int *__restrict__ neighbour, GeomP g,
                           const double *__restrict__ u, const double *__restrict__ dfo,
                           double *__restrict__ dudxi) {
  const int c = xcd_cell(ntot);
  if (c >= ntot) return;
  if (c >= n) {
    dudxi[3 * (size_t)c] = 0.0;
    dudxi[3 * (size_t)c + 1] = 0.0;
    dudxi[3 * (size_t)c + 2] = 0.0;
    return;
  }
  double gx = 0.0, gy = 0.0, gz = 0.0;
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int f = cf[k];
    if (f < Fi) {
      const int p = owner[f], q = neighbour[f];
      const double fxn = g.facint[f];
      const double fxp = 1.0 - fxn;
      const double xi = g.xc[p] * fxp + g.xc[q] * fxn;
      const double yi = g.yc[p] * fxp + g.yc[q] * fxn;
      const double zi = g.zc[p] * fxp + g.zc[q] * fxn;
      double dfxi = 0.0, dfyi = 0.0, dfzi = 0.0;
      if (dfo) {
        dfxi = dfo[3 * (size_t)p] * fxp + dfo[3 * (size_t)q] * fxn;
        dfyi = dfo[3 * (size_t)p + 1] * fxp + dfo[3 * (size_t)q + 1] * fxn;
        dfzi = dfo[3 * (size_t)p + 2] * fxp + dfo[3 * (size_t)q + 2] * fxn;
      }
      const double fie = (((u[p] * fxp + u[q] * fxn) + dfxi * (g.xf[f] - xi)) +
                          dfyi * (g.yf[f] - yi)) + dfzi * (g.zf[f] - zi);
      const double dfxe = fie * g.arx[f], dfye = fie * g.ary[f], dfze = fie * g.arz[f];
      if (c == p) {
        gx = gx + dfxe;
        gy = gy + dfye;
        gz = gz + dfze;
      } else {
        gx = gx - dfxe;
        gy = gy - dfye;
        gz = gz - dfze;
      }
    } else {
      const double fi = u[n + (f - Fi)];
      gx = gx + fi * g.arx[f];
      gy = gy + fi * g.ary[f];
      gz = gz + fi * g.arz[f];
    }
  }
  const double volr = 1.0 / g.vol[c];
  dudxi[3 * (size_t)c] = gx * volr;
  dudxi[3 * (size_t)c + 1] = gy * volr;
  dudxi[3 * (size_t)c + 2] = gz * volr;
}

// facefluxscalar's face diffusivity (scalar_fluxes.f90:384-392): an inner
// face interpolates the cell array, game = dc(P)*(1-lambda) + dc(N)*lambda;
// a boundary face takes its owner's value.
__global__ void k_face_gamma_cells(int Fi, int F, const int *__restrict__ owner,
                                   const int *__restrict__ neighbour,
                                   const double *__restrict__ facint,
                                   const double *__restrict__ dc, double *__restrict__ gf) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  if (f < Fi) {
    const double fxn = facint[f];
    const double fxp = 1.0 - fxn;
    gf[f] = dc[owner[f]] * fxp + dc[neighbour[f]] * fxn;
  } else {
    gf[f] = dc[owner[f]];
  }
}

// updateBoundary.f90: zero-gradient copy (:31-44); a periodic face and its
// twin take half the sum of their owners' values (:46-68) -- the twin unless
// its own patch, coming later in the patch loop, copies its owner's value
// (per.twb).  Only cell values are read, so every face's final value is
// computed independently.
__global__ void k_update_boundary(int n, int Fi, int Fb, const int *__restrict__ owner,
                                  const int *__restrict__ bpatch, const int *__restrict__ bctype,
                                  PerP per, double *__restrict__ phi) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= Fb) return;
  if (per.np > 0) {
    const int k = per.kof[b] >= 0 ? per.kof[b] : per.twb[b];
    if (k >= 0) {
      phi[n + b] = 0.5 * (phi[owner[per.fp[k]]] + phi[owner[per.ft[k]]]);
      return;
    }
  }
  const int t = bctype[bpatch[b]];
  if (t == FCP_BC_OUTLET || t == FCP_BC_SYMMETRY || t == FCP_BC_PRESSURE || t == FCP_BC_EMPTY)
    phi[n + b] = phi[owner[Fi + b]];
}

// The periodic pairs' entries of a matrix assembled without them (laplacian
// clears the matrix, laplacian.f90:35, and sets inner and diagonal entries).
__global__ void k_zero_periodic(int Fi, int np, const int *__restrict__ slot_pn,
                                const int *__restrict__ slot_np, double *__restrict__ a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int s1 = slot_pn[Fi + k], s2 = slot_np[Fi + k];
  if (s1 >= 0) a[s1] = 0.0;
  if (s2 >= 0) a[s2] = 0.0;
}

// Convection-diffusion assembly as calculate_temperature composes it
// (temperature.f90:118-366) with facefluxsc (scalar_fluxes.f90:32-141,
// 256-337) and face_value (interpolation.f90:28-157, 318-356).  One thread
// per cell gathers its faces in ascending face id: off-diagonals written
// through the face->slot maps, the diagonal accumulator and su/sp in the
// reference's face order, then a_ii = sp - (sum(row) - a_ii) and URF.
struct FaceFlux {
  double can, cap, suadd;
};

// facefluxsc_periodic (scalar_fluxes.f90:147-251) for periodic face fp with
// owner p and twin owner q: half weights, the distance vector twice the
// owner-to-face vector, CDS for the deferred correction, and -- as the
// reference calls it with the face's index i in its patch -- Df of inner
// face i.
__device__ __forceinline__ FaceFlux periodic_flux(int fp, int i, int p, int q, const GeomP &g,
                                                  double fm, double dc,
                                                  const double *__restrict__ phi,
                                                  const double *__restrict__ gr, double gam) {
  const double fxn = 0.5, fxp = fxn;
  const double xpn = 2 * (g.xf[fp] - g.xc[p]), ypn = 2 * (g.yf[fp] - g.yc[p]),
               zpn = 2 * (g.zf[fp] - g.zc[p]);
  const double Dfi = g.Df[i];
  const double de = dc * Dfi;
  const double ce = fm < 0.0 ? fm : 0.0;
  const double cp = fm > 0.0 ? fm : 0.0;
  FaceFlux r;
  r.can = -de + ce;
  r.cap = -de - cp;
  double dfixi = gr[3 * p] * fxp + gr[3 * q] * fxn;
  double dfiyi = gr[3 * p + 1] * fxp + gr[3 * q + 1] * fxn;
  double dfizi = gr[3 * p + 2] * fxp + gr[3 * q + 2] * fxn;
  dfixi = dfixi * (g.arx[fp] - Dfi * xpn);
  dfiyi = dfiyi * (g.ary[fp] - Dfi * ypn);
  dfizi = dfizi * (g.arz[fp] - Dfi * zpn);
  const double fdfie = dc * ((dfixi + dfiyi) + dfizi);
  const int u = fm >= 0.0 ? p : q, d = fm >= 0.0 ? q : p;
  const double fii = phi[u] + (phi[d] - phi[u]) * fxp;
  double fcfie = fm * fii;
  const double fcfii = ce * phi[q] + cp * phi[p];
  fcfie = gam * (fcfie - fcfii);
  r.suadd = -fcfie + fdfie;
  return r;
}

__device__ __forceinline__ FaceFlux inner_flux(int f, int p, int q, const GeomP &g, double fm,
                                               double dc, const double *__restrict__ phi,
                                               const double *__restrict__ gr, double gam,
                                               int scheme) {
  const double fxn = g.facint[f], fxp = 1.0 - g.facint[f];
  const double xpn = g.xc[q] - g.xc[p], ypn = g.yc[q] - g.yc[p], zpn = g.zc[q] - g.zc[p];
  const double Dff = g.Df[f];
  const double de = dc * Dff;
  const double ce = fm < 0.0 ? fm : 0.0;
  const double cp = fm > 0.0 ? fm : 0.0;
  FaceFlux r;
  r.can = -de + ce;
  r.cap = -de - cp;
  double dfixi = gr[3 * p] * fxp + gr[3 * q] * fxn;
  double dfiyi = gr[3 * p + 1] * fxp + gr[3 * q + 1] * fxn;
  double dfizi = gr[3 * p + 2] * fxp + gr[3 * q + 2] * fxn;
  dfixi = dfixi * (g.arx[f] - Dff * xpn);
  dfiyi = dfiyi * (g.ary[f] - Dff * ypn);
  dfizi = dfizi * (g.arz[f] - Dff * zpn);
  const double fdfie = dc * ((dfixi + dfiyi) + dfizi);
  // upwind cell u, downwind d, interpolation weight as passed by facefluxsc
  const int u = fm >= 0.0 ? p : q, d = fm >= 0.0 ? q : p;
  const double lam = fm >= 0.0 ? fxp : fxn;
  double fii;
  if (scheme == 0) {
    fii = phi[u] + (phi[d] - phi[u]) * lam;
  } else {
    const double gfr = (gr[3 * u] * (g.xf[f] - g.xc[u]) + gr[3 * u + 1] * (g.yf[f] - g.yc[u])) +
                       gr[3 * u + 2] * (g.zf[f] - g.zc[u]);
    fii = phi[u] + gfr;
  }
  double fcfie = fm * fii;
  const double fcfii = ce * phi[q] + cp * phi[p];
  fcfie = gam * (fcfie - fcfii);
  r.suadd = -fcfie + fdfie;
  return r;
}

// Boundary kinds: 0 Dirichlet wall, 1 zero-gradient wall, 2 inlet/outlet/
// pressure flux face, 3 no term, 4 prescribed normal gradient (wallQFlux,
// :231-262), 5 periodic face (:262-313): the pair's coefficients enter the
// rows of both owners at the periodic face's place in the boundary loop
// (the twin owner's row merges its pairs, per.tw_*, into its face order).
// srd_map (may be null) names, per boundary face, the face whose |S|/dn a
// wall term reads.  With cn the Crank-Nicolson block (:325-347)
// runs between the boundary terms and the diagonal: the row is halved, the
// old neighbour values enter su through the halved coefficients, and
// (den V/dt + halved off-diagonal sum) phi_o and den V/dt join su and sp.
__global__ void k_convdiff(int n, int Fi, const int *__restrict__ cf_ptr, const int *__restrict__ cf,
                           const int *__restrict__ owner, const int *__restrict__ neighbour, GeomP g,
                           const double *__restrict__ srdb, const double *__restrict__ flmass,
                           const double *__restrict__ dcoef, double *__restrict__ phi,
                           const double *__restrict__ gr, double gam, int scheme,
                           const unsigned char *__restrict__ bc_kind, double urf,
                           const int *__restrict__ slot_pn, const int *__restrict__ slot_np,
                           const int *__restrict__ ia, const int *__restrict__ diag,
                           double *__restrict__ a, double *__restrict__ su,
                           double *__restrict__ sp, const int *__restrict__ srd_map, int cn,
                           const double *__restrict__ den, double dt,
                           const double *__restrict__ phi_o, PerP per) {
  const int c = xcd_cell(n);
  if (c >= n) return;
  double ad = 0.0, s = su[c], spc = sp[c];
  // periodic pairs whose twin face c owns, in the order of their faces
  int tw = per.np > 0 ? per.tw_ptr[c] : 0;
  const int twe = per.np > 0 ? per.tw_ptr[c + 1] : 0;
  auto twin_side = [&](int before) {
    for (; tw < twe && per.fp[per.tw_k[tw]] < before; ++tw) {
      const int kp = per.tw_k[tw], fp = per.fp[kp];
      if (bc_kind[fp - Fi] != 5) continue;
      const FaceFlux ff = periodic_flux(fp, per.ii[kp], owner[fp], c, g, flmass[fp], dcoef[fp],
                                        phi, gr, gam);
      const int sl = slot_np[Fi + kp];
      if (sl >= 0) a[sl] = ff.cap;
      ad = ad - ff.cap;
      s = s - ff.suadd;
    }
  };
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int f = cf[k];
    if (f >= Fi) twin_side(f);
    if (f < Fi) {
      const int p = owner[f], q = neighbour[f];
      const FaceFlux ff = inner_flux(f, p, q, g, flmass[f], dcoef[f], phi, gr, gam, scheme);
      if (c == p) {
        const int sl = slot_pn[f];
        if (sl >= 0) a[sl] = ff.can;
        ad = ad - ff.can;
        s = s + ff.suadd;
      } else {
        const int sl = slot_np[f];
        if (sl >= 0) a[sl] = ff.cap;
        ad = ad - ff.cap;
        s = s - ff.suadd;
      }
    } else {
      const int b = f - Fi;
      const int kind = bc_kind[b];
      if (kind == 2) {  // facefluxsc_boundary, :182-190
        const double xpn = g.xf[f] - g.xc[c], ypn = g.yf[f] - g.yc[c], zpn = g.zf[f] - g.zc[c];
        const double ax = g.arx[f], ay = g.ary[f], az = g.arz[f];
        const double Dfi = ((ax * ax + ay * ay) + az * az) / ((xpn * ax + ypn * ay) + zpn * az);
        const double de = dcoef[f] * Dfi;
        const double fm = flmass[f];
        const double ce = fm < 0.0 ? fm : 0.0;
        const double can = -de + ce;
        const double dfixi = gr[3 * c] * (ax - Dfi * xpn);
        const double dfiyi = gr[3 * c + 1] * (ay - Dfi * ypn);
        const double dfizi = gr[3 * c + 2] * (az - Dfi * zpn);
        const double suadd = dcoef[f] * ((dfixi + dfiyi) + dfizi);
        spc = spc - can;
        s = s - can * phi[n + b] + suadd;
      } else if (kind == 0) {  // Dirichlet wall :206-210
        const double dw = dcoef[f] * srdb[srd_map ? srd_map[b] : b];
        spc = spc + dw;
        s = s + dw * phi[n + b];
      } else if (kind == 1) {  // zero-gradient wall :224
        phi[n + b] = phi[c];
      } else if (kind == 4) {  // prescribed normal gradient :231-262
        const double dw = dcoef[f] * srdb[srd_map ? srd_map[b] : b];
        const double ax = g.arx[f], ay = g.ary[f], az = g.arz[f];
        const double are = sqrt((ax * ax + ay * ay) + az * az);
        const double nxf = ax / are, nyf = ay / are, nzf = az / are;
        const size_t jb = 3 * ((size_t)n + b);
        const double dtn = (gr[jb] * nxf + gr[jb + 1] * nyf) + gr[jb + 2] * nzf;
        s = s + dw * dtn;
      } else if (kind == 5 && per.np > 0 && per.kof[b] >= 0) {  // periodic, owner side
        const int kp = per.kof[b];
        const FaceFlux ff = periodic_flux(f, per.ii[kp], c, owner[per.ft[kp]], g, flmass[f],
                                          dcoef[f], phi, gr, gam);
        const int sl = slot_pn[Fi + kp];
        if (sl >= 0) a[sl] = ff.can;
        ad = ad - ff.can;
        s = s + ff.suadd;
      }
    }
  }
  twin_side(0x7fffffff);
  const int dg = diag[c];
  a[dg] = ad;
  if (cn) {
    for (int k = ia[c]; k < ia[c + 1]; ++k) a[k] = 0.5 * a[k];
    for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
      const int f = cf[k];
      if (f >= Fi) continue;
      if (c == owner[f]) {
        const int sl = slot_pn[f];
        if (sl >= 0) s = s - a[sl] * phi_o[neighbour[f]];
      } else {
        const int sl = slot_np[f];
        if (sl >= 0) s = s - a[sl] * phi_o[owner[f]];
      }
    }
    const double apotime = den[c] * g.vol[c] / dt;
    double rh = 0.0;
    for (int k = ia[c]; k < ia[c + 1]; ++k) rh = rh + a[k];
    s = s + (apotime + (rh - a[dg])) * phi_o[c];
    spc = spc + apotime;
  }
  // off_diagonal_terms = sum(a(ia(i):ia(i+1)-1)) - a(diag(i)) (:359-360)
  double rs = 0.0;
  for (int k = ia[c]; k < ia[c + 1]; ++k) rs = rs + a[k];
  const double off = rs - a[dg];
  double aii = spc - off;
  aii = aii * (1.0 / urf);
  a[dg] = aii;
  su[c] = s + (1.0 - urf) * aii * phi[c];
  sp[c] = spc;
}

// Time terms of the transport equations (temperature.f90:82-102):
// scheme 1 = BDF1 / Crank-Nicolson volume part, 2 = BDF2.
__global__ void k_ddt(int n, int scheme, const double *__restrict__ den,
                      const double *__restrict__ vol, double dt, const double *__restrict__ fo,
                      const double *__restrict__ foo, double *__restrict__ su,
                      double *__restrict__ sp) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double apotime = den[c] * vol[c] / dt;
  if (scheme == 2) {
    su[c] = su[c] + apotime * (2.0 * fo[c] - 0.5 * foo[c]);
    sp[c] = sp[c] + 1.5 * apotime;
  } else {
    su[c] = su[c] + apotime * fo[c];
    sp[c] = sp[c] + apotime;
  }
}

// Face diffusion coefficients of calculate_temperature from the cell
// viscosity: inner faces viscos*prl + ((vis_P + (vis_N - vis_P) lambda) -
// viscos)*prtr (temperature.f90:125), flux faces viscos*prl + (vis_P -
// viscos)/sigt (:178), wall faces viscos/pranl + (vis_P - viscos)/sigt
// (:203, :237); 0 on faces without a term.
__global__ void k_transport_dcoef(int Fi, int F, const int *__restrict__ owner,
                                  const int *__restrict__ neighbour,
                                  const double *__restrict__ facint, const double *__restrict__ vis,
                                  const unsigned char *__restrict__ bc_kind, double viscos,
                                  double prl, double prtr, double pranl, double sigt, PerP per,
                                  double *__restrict__ dcoef) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const int p = owner[f];
  double d;
  if (f < Fi) {
    const int q = neighbour[f];
    d = viscos * prl + ((vis[p] + (vis[q] - vis[p]) * facint[f]) - viscos) * prtr;
  } else {
    const int kind = bc_kind[f - Fi];
    if (kind == 2)
      d = viscos * prl + (vis[p] - viscos) / sigt;
    else if (kind == 0 || kind == 4)
      d = viscos / pranl + (vis[p] - viscos) / sigt;
    else if (kind == 5 && per.np > 0 && per.kof[f - Fi] >= 0) {
      // :276, with the interpolation factor of inner face i
      const int k = per.kof[f - Fi], q = owner[per.ft[k]];
      d = viscos * prl + ((vis[p] + (vis[q] - vis[p]) * facint[per.ii[k]]) - viscos) * prtr;
    } else
      d = 0.0;
  }
  dcoef[f] = d;
}

// Smallest and largest cell value: per-block pairs, then one block.
__global__ __launch_bounds__(BLK) void k_minmax(int n, const double *__restrict__ x,
                                                double *__restrict__ part) {
  __shared__ double slo[BLK / 64], shi[BLK / 64];
  double lo = INFINITY, hi = -INFINITY;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) {
    lo = fmin(lo, x[i]);
    hi = fmax(hi, x[i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    slo[threadIdx.x >> 6] = lo;
    shi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < BLK / 64; ++w) {
      lo = fmin(lo, slo[w]);
      hi = fmax(hi, shi[w]);
    }
    part[2 * blockIdx.x] = lo;
    part[2 * blockIdx.x + 1] = hi;
  }
}

__global__ void k_minmax_final(int nb, double *__restrict__ part) {
  if (threadIdx.x != 0) return;
  double lo = part[0], hi = part[1];
  for (int b = 1; b < nb; ++b) {
    lo = fmin(lo, part[2 * b]);
    hi = fmax(hi, part[2 * b + 1]);
  }
  part[0] = lo;
  part[1] = hi;
}

// calculate_temperature's clip (:381): when the smallest value is negative,
// every cell value below `small` becomes `small`.
__global__ void k_clip_small(int n, double *__restrict__ x, const double *__restrict__ mm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !(mm[0] < 0.0)) return;
  x[i] = fmax(x[i], FCP_SMALL);
}

// ---------------------------------------------------------------------------
// Pressure correction of SIMPLE (calcp_simple.f90:69-419): incompressible,
// no pressure patches, pscheme 'linear', one corrector.
// ---------------------------------------------------------------------------
struct RcIn {
  const double *u, *v, *w, *p, *dp, *den, *apu, *apv, *apw;
};

// facefluxmass2 (faceflux_mass.f90:175-252): pressure-correction coefficient
// and Rhie-Chow mass flux of inner face f.
__device__ __forceinline__ double rc_inner(const GeomP &g, int f, int p, int q, const RcIn &in,
                                           double &fm) {
  const double lam = g.facint[f];
  const double fxn = lam, fxp = 1.0 - lam;
  const double xpn = g.xc[q] - g.xc[p], ypn = g.yc[q] - g.yc[p], zpn = g.zc[q] - g.zc[p];
  const double dene = in.den[p] * fxp + in.den[q] * fxn;
  const double Kj = g.vol[p] * in.apu[p] * fxp + g.vol[q] * in.apu[q] * fxn;
  const double cap = -(dene * Kj * g.Df[f]);
  const double ui = in.u[p] + (in.u[q] - in.u[p]) * lam;  // face_value_cds
  const double vi = in.v[p] + (in.v[q] - in.v[p]) * lam;
  const double wi = in.w[p] + (in.w[q] - in.w[p]) * lam;
  const double dpxi = (in.dp[3 * (size_t)q] * fxp + in.dp[3 * (size_t)p] * fxn) * xpn;
  const double dpyi = (in.dp[3 * (size_t)q + 1] * fxp + in.dp[3 * (size_t)p + 1] * fxn) * ypn;
  const double dpzi = (in.dp[3 * (size_t)q + 2] * fxp + in.dp[3 * (size_t)p + 2] * fxn) * zpn;
  fm = dene * ((ui * g.arx[f] + vi * g.ary[f]) + wi * g.arz[f]) +
       cap * ((((in.p[q] - in.p[p]) - dpxi) - dpyi) - dpzi);
  return cap;
}

// facefluxmass2_periodic (:313-372): periodic face f (index i in its patch,
// whose Df(i) -- an inner face's -- the reference reads), p its owner, q the
// twin face's owner, lambda = 1/2, velocities weighted by the other cell's
// 1/a_P.
__device__ __forceinline__ double rc_periodic(const GeomP &g, int f, int i, int p, int q,
                                              const RcIn &in, double &fm) {
  const double fxn = 0.5, fxp = 1.0 - 0.5;
  const double xpn = 2.0 * (g.xf[f] - g.xc[p]), ypn = 2.0 * (g.yf[f] - g.yc[p]),
               zpn = 2.0 * (g.zf[f] - g.zc[p]);
  const double dene = in.den[p] * fxp + in.den[q] * fxn;
  const double Kj0 = g.vol[p] * in.apu[p] * fxp + g.vol[q] * in.apu[q] * fxn;
  const double cap = -(dene * Kj0 * g.Df[i]);
  const double Kj = (in.apu[p] + in.apu[q]) + FCP_SMALL;
  const double ui = (in.u[p] * in.apu[q] + in.u[q] * in.apu[p]) / Kj;
  const double vi = (in.v[p] * in.apv[q] + in.v[q] * in.apv[p]) / Kj;
  const double wi = (in.w[p] * in.apw[q] + in.w[q] * in.apw[p]) / Kj;
  const double dpxi = (in.dp[3 * (size_t)q] * fxp + in.dp[3 * (size_t)p] * fxn) * xpn;
  const double dpyi = (in.dp[3 * (size_t)q + 1] * fxp + in.dp[3 * (size_t)p + 1] * fxn) * ypn;
  const double dpzi = (in.dp[3 * (size_t)q + 2] * fxp + in.dp[3 * (size_t)p + 2] * fxn) * zpn;
  fm = dene * ((ui * g.arx[f] + vi * g.ary[f]) + wi * g.arz[f]) +
       cap * ((((in.p[q] - in.p[p]) - dpxi) - dpyi) - dpzi);
  return cap;
}

__device__ __forceinline__ int btype(const int *bpatch, const int *bctype, int b) {
  return bctype[bpatch[b]];
}

// adjustMassFlow (:833-912), first pass: outlet values of u, v, w copied
// from the owner and the outlet mass fluxes from them.
__global__ void k_pc_outlet(int n, int Fi, int Fb, const int *__restrict__ owner,
                            const int *__restrict__ bpatch, const int *__restrict__ bctype,
                            GeomP g, const double *__restrict__ den, double *__restrict__ u,
                            double *__restrict__ v, double *__restrict__ w,
                            double *__restrict__ flmass) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= Fb || btype(bpatch, bctype, b) != FCP_BC_OUTLET) return;
  const int f = Fi + b, p = owner[f];
  const size_t jb = (size_t)n + b;
  const double ub = u[p], vb = v[p], wb = w[p];
  u[jb] = ub;
  v[jb] = vb;
  w[jb] = wb;
  flmass[f] = den[p] * ((ub * g.arx[f] + vb * g.ary[f]) + wb * g.arz[f]);
}

// The outlet mass flow summed in the reference's order (patches in order,
// faces in order: ascending boundary face), then fac = flomas/(flowo+small).
// One workgroup: the values of 256 faces staged in LDS, summed by one lane.
__global__ void __launch_bounds__(256) k_pc_flowo(int Fi, int Fb, const int *__restrict__ bpatch,
                                                  const int *__restrict__ bctype,
                                                  const double *__restrict__ flmass,
                                                  double flomas, double *__restrict__ fac) {
  __shared__ double val[256];
  __shared__ int use[256];
  double flowo = 0.0;
  for (int base = 0; base < Fb; base += 256) {
    const int b = base + (int)threadIdx.x;
    const bool o = b < Fb && btype(bpatch, bctype, b) == FCP_BC_OUTLET;
    use[threadIdx.x] = o ? 1 : 0;
    val[threadIdx.x] = o ? flmass[Fi + b] : 0.0;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int j = 0; j < 256; ++j)
        if (use[j]) flowo = flowo + val[j];
    __syncthreads();
  }
  if (threadIdx.x == 0) fac[0] = flomas / (flowo + FCP_SMALL);
}

__global__ void k_pc_scale(int n, int Fi, int Fb, const int *__restrict__ bpatch,
                           const int *__restrict__ bctype, const double *__restrict__ fac,
                           double *__restrict__ u, double *__restrict__ v, double *__restrict__ w,
                           double *__restrict__ flmass) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= Fb || btype(bpatch, bctype, b) != FCP_BC_OUTLET) return;
  const double fc = fac[0];
  const size_t jb = (size_t)n + b;
  flmass[Fi + b] = flmass[Fi + b] * fc;
  u[jb] = u[jb] * fc;
  v[jb] = v[jb] * fc;
  w[jb] = w[jb] * fc;
}

// The pressure-correction rows (calcp_simple.f90:78-228): every row gathers
// its faces in ascending order -- inner faces, then its boundary faces, with
// the periodic pairs whose twin face it owns merged in at their periodic
// face's place.  The owner of a face writes the face's new mass flux.
__global__ void k_pc_assemble(int n, int Fi, const int *__restrict__ cf_ptr,
                              const int *__restrict__ cf, const int *__restrict__ owner,
                              const int *__restrict__ neighbour, const int *__restrict__ bpatch,
                              const int *__restrict__ bctype, GeomP g, RcIn in,
                              const int *__restrict__ slot_pn, const int *__restrict__ slot_np,
                              const int *__restrict__ diag, double *__restrict__ flmass,
                              double *__restrict__ a, double *__restrict__ su, PerP per) {
  const int c = xcd_cell(n);
  if (c >= n) return;
  double ad = 0.0, s = 0.0;
  int tw = per.np > 0 ? per.tw_ptr[c] : 0;
  const int twe = per.np > 0 ? per.tw_ptr[c + 1] : 0;
  auto twin_side = [&](int before) {
    for (; tw < twe && per.fp[per.tw_k[tw]] < before; ++tw) {
      const int kp = per.tw_k[tw], fp = per.fp[kp];
      double fm;
      const double cap = rc_periodic(g, fp, per.ii[kp], owner[fp], c, in, fm);
      const int sl = slot_np[Fi + kp];
      if (sl >= 0) a[sl] = cap;
      ad = ad - cap;
      s = s + fm;
    }
  };
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int f = cf[k];
    if (f < Fi) {
      const int p = owner[f], q = neighbour[f];
      double fm;
      const double cap = rc_inner(g, f, p, q, in, fm);
      const int sl = (c == p) ? slot_pn[f] : slot_np[f];
      if (sl >= 0) a[sl] = cap;
      if (c == p) {
        flmass[f] = fm;
        s = s - fm;
      } else {
        s = s + fm;
      }
      ad = ad - cap;
    } else {
      twin_side(f);
      const int b = f - Fi, t = btype(bpatch, bctype, b);
      if (t == FCP_BC_INLET || t == FCP_BC_OUTLET) {
        s = s - flmass[f];
      } else if (t == FCP_BC_PERIODIC && per.np > 0 && per.kof[b] >= 0) {
        const int kp = per.kof[b];
        double fm;
        const double cap = rc_periodic(g, f, per.ii[kp], c, owner[per.ft[kp]], in, fm);
        const int sl = slot_pn[Fi + kp];
        if (sl >= 0) a[sl] = cap;
        flmass[f] = fm;
        ad = ad - cap;
        s = s - fm;
      }
    }
  }
  twin_side(0x7fffffff);
  a[diag[c]] = ad;
  su[c] = s;
}

// Mass-flux correction (:333-376): inner faces, and periodic faces whose twin
// takes the same corrected flux.
__global__ void k_pc_fluxcorr(int Fi, int F, const int *__restrict__ owner,
                              const int *__restrict__ neighbour, const int *__restrict__ slot_pn,
                              const double *__restrict__ a, const double *__restrict__ pp,
                              double *__restrict__ flmass, PerP per) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  if (f < Fi) {
    const int p = owner[f], q = neighbour[f], sl = slot_pn[f];
    if (sl >= 0) flmass[f] = flmass[f] + a[sl] * (pp[q] - pp[p]);
  } else if (per.np > 0 && per.kof[f - Fi] >= 0) {
    const int kp = per.kof[f - Fi], ft = per.ft[kp], sl = slot_pn[Fi + kp];
    const int p = owner[f], q = owner[ft];
    if (sl < 0) return;
    const double fm = flmass[f] + a[sl] * (pp[q] - pp[p]);
    flmass[f] = fm;
    flmass[ft] = fm;
  }
}

// bpres (bpres.f90): stage 1 copies the owner's value to every boundary face
// (no pressure patches here); stage 2 extrapolates to wall faces with the
// stage-1 gradient.
__global__ void k_pc_bpres(int n, int Fi, int Fb, int stage, const int *__restrict__ owner,
                           const int *__restrict__ bpatch, const int *__restrict__ bctype,
                           GeomP g, const double *__restrict__ dp, double *__restrict__ pp) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= Fb) return;
  const int f = Fi + b, p = owner[f];
  if (stage == 1) {
    pp[(size_t)n + b] = pp[p];
  } else if (btype(bpatch, bctype, b) == FCP_BC_WALL) {
    const double xpb = g.xf[f] - g.xc[p], ypb = g.yf[f] - g.yc[p], zpb = g.zf[f] - g.zc[p];
    pp[(size_t)n + b] = ((pp[p] + dp[3 * (size_t)p] * xpb) + dp[3 * (size_t)p + 1] * ypb) +
                        dp[3 * (size_t)p + 2] * zpb;
  }
}

// gradp_and_sources (nablap.f90:19-209), one bpres stage per launch: the
// inner-face sums -(sum pf S) (first launch: computed and kept in su/sv/sw;
// later: read back), the gradient (-s + sum_b pb S) / vol, and with `last`
// the boundary terms s - sum_b pb S, u += su apu, v += sv apv, w += sw apw
// and p += urfp (pp - pp[pref]) (calcp_simple.f90:398-404).
__global__ void k_pc_grad(int n, int Fi, const int *__restrict__ cf_ptr,
                          const int *__restrict__ cf, const int *__restrict__ owner,
                          const int *__restrict__ neighbour, GeomP g, const double *__restrict__ pp,
                          int first, int last, double *__restrict__ dp, double *__restrict__ su,
                          double *__restrict__ sv, double *__restrict__ sw,
                          const double *__restrict__ apu, const double *__restrict__ apv,
                          const double *__restrict__ apw, double urfp, int pref,
                          double *__restrict__ u, double *__restrict__ v, double *__restrict__ w,
                          double *__restrict__ pr) {
  const int c = xcd_cell(n);
  if (c >= n) return;
  double sx, sy, sz;
  const int k0 = cf_ptr[c], k1 = cf_ptr[c + 1];
  if (first) {
    sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = k0; k < k1; ++k) {
      const int f = cf[k];
      if (f >= Fi) break;
      const int p = owner[f], q = neighbour[f];
      const double pf = pp[p] + (pp[q] - pp[p]) * g.facint[f];
      const double dfxe = pf * g.arx[f], dfye = pf * g.ary[f], dfze = pf * g.arz[f];
      if (c == p) {
        sx = sx - dfxe;
        sy = sy - dfye;
        sz = sz - dfze;
      } else {
        sx = sx + dfxe;
        sy = sy + dfye;
        sz = sz + dfze;
      }
    }
    su[c] = sx;
    sv[c] = sy;
    sw[c] = sz;
  } else {
    sx = su[c];
    sy = sv[c];
    sz = sw[c];
  }
  double dx = -sx, dy = -sy, dz = -sz;
  for (int k = k0; k < k1; ++k) {
    const int f = cf[k];
    if (f < Fi) continue;
    const double pb = pp[(size_t)n + (f - Fi)];
    dx = dx + pb * g.arx[f];
    dy = dy + pb * g.ary[f];
    dz = dz + pb * g.arz[f];
  }
  const double volr = 1.0 / g.vol[c];
  dp[3 * (size_t)c] = dx * volr;
  dp[3 * (size_t)c + 1] = dy * volr;
  dp[3 * (size_t)c + 2] = dz * volr;
  if (!last) return;
  for (int k = k0; k < k1; ++k) {
    const int f = cf[k];
    if (f < Fi) continue;
    const double pb = pp[(size_t)n + (f - Fi)];
    sx = sx - pb * g.arx[f];
    sy = sy - pb * g.ary[f];
    sz = sz - pb * g.arz[f];
  }
  su[c] = sx;
  sv[c] = sy;
  sw[c] = sz;
  if (!apu) return;  // gradp_and_sources alone (the momentum sources)
  u[c] = u[c] + sx * apu[c];
  v[c] = v[c] + sy * apv[c];
  w[c] = w[c] + sz * apw[c];
  pr[c] = pr[c] + urfp * (pp[c] - pp[pref]);
}

// updateVelocityAtBoundary (velocity.f90:1184-1279): empty and periodic faces
// take the owner's velocity, symmetry faces its reflection with the
// (unnormalised) area vector.
__global__ void k_pc_velbnd(int n, int Fi, int Fb, const int *__restrict__ owner,
                            const int *__restrict__ bpatch, const int *__restrict__ bctype,
                            GeomP g, double *__restrict__ u, double *__restrict__ v,
                            double *__restrict__ w) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= Fb) return;
  const int t = btype(bpatch, bctype, b);
  const int f = Fi + b, p = owner[f];
  const size_t jb = (size_t)n + b;
  if (t == FCP_BC_EMPTY || t == FCP_BC_PERIODIC) {
    u[jb] = u[p];
    v[jb] = v[p];
    w[jb] = w[p];
  } else if (t == FCP_BC_SYMMETRY) {
    const double Unmag = (u[p] * g.arx[f] + v[p] * g.ary[f]) + w[p] * g.arz[f];
    u[jb] = u[p] - Unmag * g.arx[f];
    v[jb] = v[p] - Unmag * g.ary[f];
    w[jb] = w[p] - Unmag * g.arz[f];
  }
}

// ---------------------------------------------------------------------------
// Momentum assembly of calcuvw (velocity.f90:50-750) before its three solves.
// ---------------------------------------------------------------------------
struct MomIn {
  const double *U[3], *G[3], *O[3], *OO[3];
  const double *den, *vis, *visw, *flmass;
  const int *wall_of;  // [Fb] index of the face among the wall faces in patch order, or -1
  double viscos, dt, gds, urf[3];
  int ts, scheme;
};

// sngrad (gradients.f90:1720-1781) of component phi at inner face f.
__device__ __forceinline__ void mom_sngrad(const GeomP &g, int f, int p, int q,
                                           const double *__restrict__ phi,
                                           const double *__restrict__ gr, double di[3],
                                           double dii[3]) {
  const double lam = g.facint[f];
  const double fxn = lam, fxp = 1.0 - lam;
  const double xpn = g.xc[q] - g.xc[p], ypn = g.yc[q] - g.yc[p], zpn = g.zc[q] - g.zc[p];
  const double ax = g.arx[f], ay = g.ary[f], az = g.arz[f];
  const double vole = (xpn * ax + ypn * ay) + zpn * az;
  const double dx = gr[3 * (size_t)p] * fxp + gr[3 * (size_t)q] * fxn;
  const double dy = gr[3 * (size_t)p + 1] * fxp + gr[3 * (size_t)q + 1] * fxn;
  const double dz = gr[3 * (size_t)p + 2] * fxp + gr[3 * (size_t)q + 2] * fxn;
  const double e = (((phi[q] - phi[p]) - dx * xpn) - dy * ypn) - dz * zpn;
  dii[0] = dx + ax / vole * e;
  dii[1] = dy + ay / vole * e;
  dii[2] = dz + az / vole * e;
  const double Dff = g.Df[f];
  di[0] = dx * (ax - Dff * xpn);
  di[1] = dy * (ay - Dff * ypn);
  di[2] = dz * (az - Dff * zpn);
}

// facefluxuvw (velocity.f90:754-878): coefficients and the three sources.
__device__ __forceinline__ void mom_inner(const GeomP &g, int f, int p, int q, const MomIn &in,
                                          double &can, double &cap, double sadd[3]) {
  const double lam = g.facint[f], fxn = lam, fxp = 1.0 - lam;
  const double fm = in.flmass[f];
  const double game = in.vis[p] + (in.vis[q] - in.vis[p]) * lam;
  const double de = game * g.Df[f];
  const double ce = fm < 0.0 ? fm : 0.0, cp = fm > 0.0 ? fm : 0.0;
  can = -de + ce;
  cap = -de - cp;
  double di[3][3], dii[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r) mom_sngrad(g, f, p, q, in.U[r], in.G[r], di[r], dii[r]);
  const double ax = g.arx[f], ay = g.ary[f], az = g.arz[f];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double fde = game * ((dii[0][r] * ax + dii[1][r] * ay) + dii[2][r] * az);
    const double fdi = game * ((di[r][0] + di[r][1]) + di[r][2]);
    fde = fde + fdi;
    const double *phi = in.U[r], *gr = in.G[r];
    const double fuds = cp * phi[p] + ce * phi[q];
    double ue;
    const int up = fm >= 0.0 ? p : q, dn = fm >= 0.0 ? q : p;
    if (in.scheme == 0) {  // face_value_cds with the weight the caller passes
      ue = phi[up] + (phi[dn] - phi[up]) * (fm >= 0.0 ? fxp : fxn);
    } else {  // face_value_2nd_upwind
      const double gfr = (gr[3 * (size_t)up] * (g.xf[f] - g.xc[up]) +
                          gr[3 * (size_t)up + 1] * (g.yf[f] - g.yc[up])) +
                         gr[3 * (size_t)up + 2] * (g.zf[f] - g.zc[up]);
      ue = phi[up] + gfr;
    }
    const double fhigh = fm * ue;
    sadd[r] = -(in.gds * (fhigh - fuds)) + fde;
  }
}

// facefluxuvw_periodic (velocity.f90:1038-1179): periodic face f, owner p,
// twin owner q.
__device__ __forceinline__ void mom_periodic(const GeomP &g, int f, int p, int q,
                                             const MomIn &in, double &can, double &cap,
                                             double sadd[3]) {
  const double fxn = 0.5, fxp = fxn;
  const double ax = g.arx[f], ay = g.ary[f], az = g.arz[f];
  const double xpn = 2.0 * (g.xf[f] - g.xc[p]), ypn = 2.0 * (g.yf[f] - g.yc[p]),
               zpn = 2.0 * (g.zf[f] - g.zc[p]);
  const double dpn = sqrt((xpn * xpn + ypn * ypn) + zpn * zpn);
  const double are = sqrt((ax * ax + ay * ay) + az * az);
  const double game = in.vis[p] * fxp + in.vis[q] * fxn;
  const double de = game * ((ax * ax + ay * ay) + az * az) / ((xpn * ax + ypn * ay) + zpn * az);
  const double fm = in.flmass[f];
  const double ce = fm < 0.0 ? fm : 0.0, cp = fm > 0.0 ? fm : 0.0;
  can = -de + ce;
  cap = -de - cp;
  double d[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      d[r][k] = in.G[r][3 * (size_t)p + k] * fxp + in.G[r][3 * (size_t)q + k] * fxn;
  const double fde[3] = {
      game * (((d[0][0] + d[0][0]) * ax + (d[0][1] + d[1][0]) * ay) + (d[0][2] + d[2][0]) * az),
      game * (((d[0][1] + d[1][0]) * ax + (d[1][1] + d[1][1]) * ay) + (d[1][2] + d[2][1]) * az),
      game * (((d[0][2] + d[2][0]) * ax + (d[2][1] + d[1][2]) * ay) + (d[2][2] + d[2][2]) * az)};
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double *phi = in.U[r];
    const double fdi = game * are / dpn * ((d[r][0] * xpn + d[r][1] * ypn) + d[r][2] * zpn);
    const double fuds = cp * phi[p] + ce * phi[q];
    const double ue = fm >= 0.0 ? phi[p] + (phi[q] - phi[p]) * fxp
                                : phi[q] + (phi[p] - phi[q]) * fxn;
    const double fhigh = fm * ue;
    sadd[r] = (-(in.gds * (fhigh - fuds)) + fde[r]) - fdi;
  }
}

// One thread per cell: the pressure sources from gradp_and_sources in s3,
// then the time terms, the cell's faces in ascending order (inner faces,
// boundary faces by type, the periodic pairs whose twin face it owns at their
// periodic face's place), the Crank-Nicolson block and the three diagonals.
// Only this thread reads or writes the entries of row c.
__global__ void k_mom_cells(int n, int Fi, const int *__restrict__ cf_ptr,
                            const int *__restrict__ cf, const int *__restrict__ owner,
                            const int *__restrict__ neighbour, const int *__restrict__ bpatch,
                            const int *__restrict__ bctype, GeomP g, MomIn in,
                            const int *__restrict__ slot_pn, const int *__restrict__ slot_np,
                            const int *__restrict__ ia, const int *__restrict__ diag,
                            double *__restrict__ a, const double *__restrict__ s3,
                            double *__restrict__ rhs3, double *__restrict__ diag3,
                            double *__restrict__ ap3, double *__restrict__ sp3,
                            double *__restrict__ src3, PerP per) {
  const int c = xcd_cell(n);
  if (c >= n) return;
  double S[3], SP[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int r = 0; r < 3; ++r) S[r] = s3[r * (size_t)n + c];
  if (in.ts > 0) {  // :226-279
    const double apotime = in.den[c] * g.vol[c] / in.dt;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      if (in.ts == 1 || in.ts == 3) {
        S[r] = S[r] + apotime * in.O[r][c];
        SP[r] = SP[r] + apotime;
      } else {
        S[r] = S[r] + apotime * (2.0 * in.O[r][c] - 0.5 * in.OO[r][c]);
        SP[r] = SP[r] + 1.5 * apotime;
      }
    }
  }
  int tw = per.np > 0 ? per.tw_ptr[c] : 0;
  const int twe = per.np > 0 ? per.tw_ptr[c + 1] : 0;
  auto twin_side = [&](int before) {
    for (; tw < twe && per.fp[per.tw_k[tw]] < before; ++tw) {
      const int kp = per.tw_k[tw], fp = per.fp[kp];
      double can, cap, sadd[3];
      mom_periodic(g, fp, owner[fp], c, in, can, cap, sadd);
      const int sl = slot_np[Fi + kp];
      if (sl >= 0) a[sl] = cap;
#pragma unroll
      for (int r = 0; r < 3; ++r) S[r] = S[r] - sadd[r];
    }
  };
  const int k0 = cf_ptr[c], k1 = cf_ptr[c + 1];
  for (int k = k0; k < k1; ++k) {
    const int f = cf[k];
    if (f < Fi) {
      const int p = owner[f], q = neighbour[f];
      double can, cap, sadd[3];
      mom_inner(g, f, p, q, in, can, cap, sadd);
      if (c == p) {
        const int sl = slot_pn[f];
        if (sl >= 0) a[sl] = can;
#pragma unroll
        for (int r = 0; r < 3; ++r) S[r] = S[r] + sadd[r];
      } else {
        const int sl = slot_np[f];
        if (sl >= 0) a[sl] = cap;
#pragma unroll
        for (int r = 0; r < 3; ++r) S[r] = S[r] - sadd[r];
      }
      continue;
    }
    twin_side(f);
    const int b = f - Fi, t = bctype[bpatch[b]];
    const size_t jb = (size_t)n + b;
    const double ax = g.arx[f], ay = g.ary[f], az = g.arz[f];
    if (t == FCP_BC_INLET || t == FCP_BC_OUTLET || t == FCP_BC_PRESSURE) {
      // facefluxuvw_bnd (velocity.f90:882-1034), scatter :328-352
      const double xpn = g.xf[f] - g.xc[c], ypn = g.yf[f] - g.yc[c], zpn = g.zf[f] - g.zc[c];
      const double vole = (xpn * ax + ypn * ay) + zpn * az;
      const double Dfi = ((ax * ax + ay * ay) + az * az) / vole;
      const double game = in.vis[jb];
      const double de = game * Dfi;
      const double fm = in.flmass[f];
      const double cb = -de + (fm < 0.0 ? fm : 0.0);
      double di[3][3], dii[3][3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double *gr = in.G[r], *phi = in.U[r];
        const double dx = gr[3 * (size_t)c], dy = gr[3 * (size_t)c + 1], dz = gr[3 * (size_t)c + 2];
        const double e = (((phi[jb] - phi[c]) - dx * xpn) - dy * ypn) - dz * zpn;
        dii[r][0] = dx + ax / vole * e;
        dii[r][1] = dy + ay / vole * e;
        dii[r][2] = dz + az / vole * e;
        di[r][0] = dx * (ax - Dfi * xpn);
        di[r][1] = dy * (ay - Dfi * ypn);
        di[r][2] = dz * (az - Dfi * zpn);
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double fde = game * ((dii[0][r] * ax + dii[1][r] * ay) + dii[2][r] * az);
        const double fdi = game * ((di[r][0] + di[r][1]) + di[r][2]);
        SP[r] = SP[r] - cb;
        S[r] = S[r] - cb * in.U[r][jb] + (fde + fdi);
      }
    } else if (t == FCP_BC_SYMMETRY) {  // :354-391, vis(inp) with inp = numCells+1
      const double are = sqrt((ax * ax + ay * ay) + az * az);
      const double arer = 1.0 / are;
      const double nf[3] = {ax * arer, ay * arer, az * arer};
      const double dpb = ((g.xf[f] - g.xc[c]) * nf[0] + (g.yf[f] - g.yc[c]) * nf[1]) +
                         (g.zf[f] - g.zc[c]) * nf[2];
      const double cf_ = 2.0 * in.vis[n] * are / dpb;
      const double uc = in.U[0][c], vc = in.U[1][c], wc = in.U[2][c];
      S[0] = S[0] - cf_ * nf[0] * (nf[1] * vc + nf[2] * wc);
      S[1] = S[1] - cf_ * nf[1] * (nf[0] * uc + nf[2] * wc);
      S[2] = S[2] - cf_ * nf[2] * (nf[0] * uc + nf[1] * vc);
#pragma unroll
      for (int r = 0; r < 3; ++r) SP[r] = SP[r] + cf_ * (nf[r] * nf[r]);
    } else if (t == FCP_BC_PERIODIC && per.np > 0 && per.kof[b] >= 0) {
      const int kp = per.kof[b];
      double can, cap, sadd[3];
      mom_periodic(g, f, c, owner[per.ft[kp]], in, can, cap, sadd);
      const int sl = slot_pn[Fi + kp];
      if (sl >= 0) a[sl] = can;
#pragma unroll
      for (int r = 0; r < 3; ++r) S[r] = S[r] + sadd[r];
    } else if (t == FCP_BC_WALL) {  // :437-482
      const double vw = in.visw[in.wall_of[b]];
      const double viss = in.viscos > vw ? in.viscos : vw;
      const double are = sqrt((ax * ax + ay * ay) + az * az);
      const double arer = 1.0 / are;
      const double nxf = ax * arer, nyf = ay * arer, nzf = az * arer;
      const double dpb = ((g.xf[f] - g.xc[c]) * nxf + (g.yf[f] - g.yc[c]) * nyf) +
                         (g.zf[f] - g.zc[c]) * nzf;
      const double vsol = viss * are / dpb;
      const double ub = in.U[0][jb], vb = in.U[1][jb], wb = in.U[2][jb];
      const double upb = in.U[0][c] - ub, vpb = in.U[1][c] - vb, wpb = in.U[2][c] - wb;
      SP[0] = SP[0] + vsol * (1.0 - nxf * nxf);
      SP[1] = SP[1] + vsol * (1.0 - nyf * nyf);
      SP[2] = SP[2] + vsol * (1.0 - nzf * nzf);
      S[0] = S[0] + vsol * ((ub * (1.0 - nxf * nxf) + vpb * nyf * nxf) + wpb * nzf * nxf);
      S[1] = S[1] + vsol * ((upb * nxf * nyf + vb * (1.0 - nyf * nyf)) + wpb * nzf * nyf);
      S[2] = S[2] + vsol * ((upb * nxf * nzf + vpb * nyf * nzf) + wb * (1.0 - nzf * nzf));
    }
  }
  twin_side(0x7fffffff);
  // Crank-Nicolson (:582-606) and the diagonals (:608-746), component by
  // component; the row sums read whatever the diagonal entry holds
  const int r0 = ia[c], r1 = ia[c + 1], dg = diag[c];
  if (in.ts == 3)
    for (int k = r0; k < r1; ++k) a[k] = 0.5 * a[k];
  for (int r = 0; r < 3; ++r) {
    if (in.ts == 3) {
      for (int k = k0; k < k1; ++k) {
        const int f = cf[k];
        if (f >= Fi) break;
        const int p = owner[f], q = neighbour[f];
        if (c == p) S[r] = S[r] - a[slot_pn[f]] * in.O[r][q];
        else S[r] = S[r] - a[slot_np[f]] * in.O[r][p];
      }
      const double apotime = in.den[c] * g.vol[c] / in.dt;
      double sum = 0.0;
      for (int k = r0; k < r1; ++k) sum = sum + a[k];
      const double off = sum - a[dg];
      S[r] = S[r] + (apotime + off) * in.O[r][c];
      SP[r] = SP[r] + apotime;
    }
    if (r > 0) a[dg] = 0.0;
    const double urfr = 1.0 / in.urf[r], urfm = 1.0 - in.urf[r];
    double sum = 0.0;
    for (int k = r0; k < r1; ++k) sum = sum + a[k];
    const double off = sum - a[dg];
    double ad = SP[r] - off;
    ap3[r * (size_t)n + c] = 1.0 / (ad + FCP_SMALL);
    ad = ad * urfr;
    a[dg] = ad;
    diag3[r * (size_t)n + c] = ad;
    rhs3[r * (size_t)n + c] = S[r] + urfm * ad * in.U[r][c];
    sp3[r * (size_t)n + c] = SP[r];
    src3[r * (size_t)n + c] = S[r];
  }
}

// ---------------------------------------------------------------------------
// Block forms of the face Laplacian, Gauss gradient and non-orthogonal
// correction: a workgroup takes 256 consecutive cells; first every inner face
// its cells own and every face reaching them from another block is computed
// once into LDS (coalesced runs of faces), then each cell sums its faces'
// terms from LDS in ascending face order, exactly as the per-cell kernels do.
// ---------------------------------------------------------------------------
constexpr int ASM_B = 256;        // cells per block
constexpr int ASM_MAXF = 1536;    // faces staged per block
constexpr int ASM_MAXE = 2560;    // CSR entries of a block's rows staged for the Laplacian

struct AsmBlk {
  const int *fo, *incp, *inc, *slot;
  int nb;
};
// Per cell-face incidence: bit 31 a boundary face (read the face id from the
// cell's face list), bit 30 the cell is the face's neighbour, bits 0-10 the
// face's LDS slot, bits 11-22 the row entry's position in the block's CSR
// range (4095: none).
constexpr int ASM_BND = (int)0x80000000u, ASM_NB = 0x40000000;
__device__ __forceinline__ int asm_j(int sl) { return sl & 0x7FF; }
__device__ __forceinline__ int asm_e(int sl) { return (sl >> 11) & 0xFFF; }

__device__ __forceinline__ int asm_block(int nb) {
  const int chunk = (nb + 7) / 8;
  return (int)(blockIdx.x & 7) * chunk + (int)(blockIdx.x >> 3);
}
int asm_grid(int nb) { return nb > 0 ? 8 * ((nb + 7) / 8) : 1; }

// face of staged position j of block b
__device__ __forceinline__ int asm_face(const AsmBlk &B, int b, int j, int nown) {
  return j < nown ? B.fo[b] + j : B.inc[B.incp[b] + (j - nown)];
}

__global__ void __launch_bounds__(ASM_B)
    k_grad_blk(int n, int Fi, AsmBlk B, const int *__restrict__ cf_ptr, const int *__restrict__ cf,
               const int *__restrict__ owner, const int *__restrict__ neighbour, GeomP g,
               const double *__restrict__ u, double *__restrict__ dudxi) {
  __shared__ double cont[3 * ASM_MAXF];
  const int b = asm_block(B.nb);
  if (b >= B.nb) return;
  const int nown = B.fo[b + 1] - B.fo[b], nf = nown + (B.incp[b + 1] - B.incp[b]);
  for (int j = threadIdx.x; j < nf; j += ASM_B) {
    const int f = asm_face(B, b, j, nown);
    const int p = owner[f], q = neighbour[f];
    const double fie = u[p] + (u[q] - u[p]) * g.facint[f];
    cont[3 * j] = fie * g.arx[f];
    cont[3 * j + 1] = fie * g.ary[f];
    cont[3 * j + 2] = fie * g.arz[f];
  }
  __syncthreads();
  const int c = b * ASM_B + (int)threadIdx.x;
  if (c >= n) return;
  double gx = 0.0, gy = 0.0, gz = 0.0;
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int sl = B.slot[k];
    if (sl >= 0) {
      const int j = asm_j(sl);
      if (!(sl & ASM_NB)) {
        gx = gx + cont[3 * j];
        gy = gy + cont[3 * j + 1];
        gz = gz + cont[3 * j + 2];
      } else {
        gx = gx - cont[3 * j];
        gy = gy - cont[3 * j + 1];
        gz = gz - cont[3 * j + 2];
      }
    } else {  // gradbc :1696-1715
      const int f = cf[k];
      const double fi = u[n + (f - Fi)];
      gx = gx + fi * g.arx[f];
      gy = gy + fi * g.ary[f];
      gz = gz + fi * g.arz[f];
    }
  }
  const double volr = 1.0 / g.vol[c];
  dudxi[3 * (size_t)c] = gx * volr;
  dudxi[3 * (size_t)c + 1] = gy * volr;
  dudxi[3 * (size_t)c + 2] = gz * volr;
}

__global__ void k_zero_rows3(int n, int ntot, double *__restrict__ x) {
  const int i = n + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntot) return;
  x[3 * (size_t)i] = 0.0;
  x[3 * (size_t)i + 1] = 0.0;
  x[3 * (size_t)i + 2] = 0.0;
}

__global__ void __launch_bounds__(ASM_B)
    k_nonorth_blk(int n, int Fi, AsmBlk B, const int *__restrict__ cf_ptr,
                  const int *__restrict__ cf, const int *__restrict__ owner,
                  const int *__restrict__ neighbour, GeomP g, const double *__restrict__ gam,
                  const double *__restrict__ gr, double *__restrict__ su) {
  __shared__ double cont[ASM_MAXF];
  const int b = asm_block(B.nb);
  if (b >= B.nb) return;
  const int nown = B.fo[b + 1] - B.fo[b], nf = nown + (B.incp[b + 1] - B.incp[b]);
  for (int j = threadIdx.x; j < nf; j += ASM_B) {
    const int f = asm_face(B, b, j, nown);
    const int p = owner[f], q = neighbour[f];
    const double fxn = g.facint[f], fxp = 1.0 - g.facint[f];
    const double xpn = g.xc[q] - g.xc[p], ypn = g.yc[q] - g.yc[p], zpn = g.zc[q] - g.zc[p];
    const double Dff = g.Df[f];
    double dfixi = gr[3 * p] * fxp + gr[3 * q] * fxn;
    double dfiyi = gr[3 * p + 1] * fxp + gr[3 * q + 1] * fxn;
    double dfizi = gr[3 * p + 2] * fxp + gr[3 * q + 2] * fxn;
    dfixi = dfixi * (g.arx[f] - Dff * xpn);
    dfiyi = dfiyi * (g.ary[f] - Dff * ypn);
    dfizi = dfizi * (g.arz[f] - Dff * zpn);
    cont[j] = gam[f] * ((dfixi + dfiyi) + dfizi);
  }
  __syncthreads();
  const int c = b * ASM_B + (int)threadIdx.x;
  if (c >= n) return;
  double s = su[c];
  for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
    const int sl = B.slot[k];
    if (sl < 0) continue;
    const int j = asm_j(sl);
    s = (sl & ASM_NB) ? s - cont[j] : s + cont[j];
  }
  su[c] = s;
}

// Laplacian rows of a block, written to a from LDS in one coalesced pass
// (entries no face writes -- the periodic pairs' -- are written as 0).
__global__ void __launch_bounds__(ASM_B)
    k_lap_blk(int n, int Fi, AsmBlk B, const int *__restrict__ cf_ptr, const int *__restrict__ cf,
              const int *__restrict__ owner, const int *__restrict__ neighbour, GeomP g,
              const double *__restrict__ mu, const double *__restrict__ mu_f,
              const unsigned char *__restrict__ bc_kind, const double *__restrict__ phi,
              const int *__restrict__ slot_pn, const int *__restrict__ slot_np,
              const int *__restrict__ ia, const int *__restrict__ diag, double *__restrict__ a,
              double *__restrict__ su) {
  __shared__ double cont[ASM_MAXF];
  __shared__ double rb[ASM_MAXE];
  const int b = asm_block(B.nb);
  if (b >= B.nb) return;
  const int c0 = b * ASM_B, c1 = min(n, c0 + ASM_B);
  const int e0 = ia[c0], ne = ia[c1] - e0;
  for (int e = threadIdx.x; e < ne; e += ASM_B) rb[e] = 0.0;
  const int nown = B.fo[b + 1] - B.fo[b], nf = nown + (B.incp[b + 1] - B.incp[b]);
  for (int j = threadIdx.x; j < nf; j += ASM_B) {
    const int f = asm_face(B, b, j, nown);
    double cap;
    if (mu_f) {
      cap = mu_f[f] * g.Df[f];
    } else {  // faceFluxLaplacian :113-151
      const int p = owner[f], q = neighbour[f];
      const double fxn = g.facint[f], fxp = 1.0 - g.facint[f];
      const double xpn = g.xc[q] - g.xc[p], ypn = g.yc[q] - g.yc[p], zpn = g.zc[q] - g.zc[p];
      const double ax = g.arx[f], ay = g.ary[f], az = g.arz[f];
      const double smdpn = ((ax * ax + ay * ay) + az * az) / ((ax * xpn + ay * ypn) + az * zpn);
      cap = (fxp * mu[p] + fxn * mu[q]) * smdpn;
    }
    cont[j] = cap;
  }
  __syncthreads();
  const int c = c0 + (int)threadIdx.x;
  if (c < n) {
    double ad = 0.0;  // a(diag) starts from the cleared matrix (:35)
    double s = su[c];
    for (int k = cf_ptr[c]; k < cf_ptr[c + 1]; ++k) {
      const int sl = B.slot[k];
      if (sl >= 0) {
        const double cap = cont[asm_j(sl)];
        const int e = asm_e(sl);
        if (e != 0xFFF) rb[e] = cap;
        ad = ad - cap;
      } else {
        const int f = cf[k];
        if (bc_kind && bc_kind[f - Fi] != 0) continue;
        const double ax = g.arx[f], ay = g.ary[f], az = g.arz[f];
        const double are = sqrt((ax * ax + ay * ay) + az * az);
        const double nxf = ax / are, nyf = ay / are, nzf = az / are;
        const double dfn = ((g.xf[f] - g.xc[c]) * nxf + (g.yf[f] - g.yc[c]) * nyf) +
                           (g.zf[f] - g.zc[c]) * nzf;
        const double dcoef = (mu_f ? mu_f[f] : mu[c]) * are / dfn;
        ad = ad - dcoef;
        s = s - dcoef * phi[n + (f - Fi)];
      }
    }
    rb[diag[c] - e0] = ad;
    su[c] = s;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ne; e += ASM_B) a[e0 + e] = rb[e];
}

GeomP geom_of(fcp_ctx *ctx) {
  return GeomP{ctx->xc.p, ctx->yc.p, ctx->zc.p, ctx->vol.p, ctx->arx.p, ctx->ary.p, ctx->arz.p,
               ctx->xf.p, ctx->yf.p, ctx->zf.p, ctx->facint.p, ctx->Df.p};
}

}  // namespace

// Cell blocks of the per-cell assembly kernels (once per CSR pattern): needs
// the inner faces sorted by owner and every block's staged faces within LDS.
static int asm_prepare(fcp_ctx *ctx) {
  if (ctx->asm_state != 0 || !ctx->have_csr) return 0;
  ctx->asm_state = -1;
  const int n = ctx->n, Fi = ctx->Fi;
  if (n == 0 || Fi == 0) return 0;
  std::vector<int> own(Fi), nbr(Fi), cfp(n + 1), ia(n + 1), spn(Fi), snp(Fi);
  FCP_HIP(ctx, hipMemcpy(own.data(), ctx->owner.p, sizeof(int) * Fi, hipMemcpyDeviceToHost));
  FCP_HIP(ctx, hipMemcpy(nbr.data(), ctx->neighbour.p, sizeof(int) * Fi, hipMemcpyDeviceToHost));
  FCP_HIP(ctx, hipMemcpy(cfp.data(), ctx->cf_ptr.p, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  FCP_HIP(ctx, hipMemcpy(ia.data(), ctx->ia.p, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  FCP_HIP(ctx, hipMemcpy(spn.data(), ctx->slot_pn.p, sizeof(int) * Fi, hipMemcpyDeviceToHost));
  FCP_HIP(ctx, hipMemcpy(snp.data(), ctx->slot_np.p, sizeof(int) * Fi, hipMemcpyDeviceToHost));
  for (int f = 1; f < Fi; ++f)
    if (own[f] < own[f - 1]) return 0;
  const int nb = (n + ASM_B - 1) / ASM_B;
  std::vector<int> fo(nb + 1), incp(nb + 1, 0);
  for (int b = 0; b <= nb; ++b)
    fo[b] = (int)(std::lower_bound(own.begin(), own.end(), b * ASM_B) - own.begin());
  for (int f = 0; f < Fi; ++f)
    if (nbr[f] / ASM_B != own[f] / ASM_B) incp[nbr[f] / ASM_B + 1]++;
  for (int b = 0; b < nb; ++b) incp[b + 1] += incp[b];
  std::vector<int> inc(incp[nb] > 0 ? incp[nb] : 1), fill(incp.begin(), incp.end() - 1);
  for (int f = 0; f < Fi; ++f)
    if (nbr[f] / ASM_B != own[f] / ASM_B) inc[fill[nbr[f] / ASM_B]++] = f;
  bool rows = true;
  for (int b = 0; b < nb; ++b) {
    if ((fo[b + 1] - fo[b]) + (incp[b + 1] - incp[b]) > ASM_MAXF) return 0;
    const int c0 = b * ASM_B, c1 = std::min(n, c0 + ASM_B);
    if (ia[c1] - ia[c0] > ASM_MAXE) rows = false;
  }
  const int ncf = cfp[n];
  std::vector<int> cf(ncf > 0 ? ncf : 1), slot(ncf > 0 ? ncf : 1, ASM_BND);
  FCP_HIP(ctx, hipMemcpy(cf.data(), ctx->cf_face.p, sizeof(int) * ncf, hipMemcpyDeviceToHost));
  for (int c = 0; c < n; ++c) {
    const int b = c / ASM_B, nown = fo[b + 1] - fo[b], e0 = ia[b * ASM_B];
    for (int k = cfp[c]; k < cfp[c + 1]; ++k) {
      const int f = cf[k];
      if (f >= Fi) continue;  // boundary: ASM_BND
      int j;
      if (f >= fo[b] && f < fo[b + 1]) {
        j = f - fo[b];
      } else {
        const int *lo = inc.data() + incp[b], *hi = inc.data() + incp[b + 1];
        const int *it = std::lower_bound(lo, hi, f);
        if (it == hi || *it != f) return 0;
        j = nown + (int)(it - lo);
      }
      const int rs = own[f] != c ? snp[f] : spn[f];
      int e = rs >= 0 ? rs - e0 : 0xFFF;
      if (e < 0 || e > 0xFFF) e = 0xFFF;
      if (rs >= 0 && e == 0xFFF) rows = false;  // outside the staged range
      slot[k] = j | (e << 11) | (own[f] != c ? ASM_NB : 0);
    }
  }
  FCP_HIP(ctx, ctx->asm_fo.alloc(nb + 1));
  FCP_HIP(ctx, ctx->asm_incp.alloc(nb + 1));
  FCP_HIP(ctx, ctx->asm_inc.alloc(inc.size()));
  FCP_HIP(ctx, ctx->asm_slot.alloc(slot.size()));
  FCP_HIP(ctx, hipMemcpy(ctx->asm_fo.p, fo.data(), sizeof(int) * fo.size(), hipMemcpyHostToDevice));
  FCP_HIP(ctx, hipMemcpy(ctx->asm_incp.p, incp.data(), sizeof(int) * incp.size(),
                         hipMemcpyHostToDevice));
  FCP_HIP(ctx, hipMemcpy(ctx->asm_inc.p, inc.data(), sizeof(int) * inc.size(),
                         hipMemcpyHostToDevice));
  FCP_HIP(ctx, hipMemcpy(ctx->asm_slot.p, slot.data(), sizeof(int) * slot.size(),
                         hipMemcpyHostToDevice));
  ctx->asm_nb = nb;
  ctx->asm_rows = rows;
  ctx->asm_state = 1;
  return 0;
}

static bool asm_on(fcp_ctx *ctx) {
  if (!ctx->tune_asm_blk) return false;
  if (asm_prepare(ctx) != 0) {
    ctx->asm_state = -1;
    return false;
  }
  return ctx->asm_state == 1;
}

static AsmBlk asm_view(const fcp_ctx *ctx) {
  return AsmBlk{ctx->asm_fo.p, ctx->asm_incp.p, ctx->asm_inc.p, ctx->asm_slot.p, ctx->asm_nb};
}

int fcp_launch_laplacian(fcp_ctx *ctx, const double *mu, const double *mu_f,
                         const unsigned char *bc_kind, const double *phi, double *su) {
  FCP_TRY(fcp_coarse_join(ctx));
  const bool blk = asm_on(ctx) && ctx->asm_rows;
  ProfScope ps(ctx, mu_f ? "laplacian_face" : "laplacian");
  if (blk)
    k_lap_blk<<<asm_grid(ctx->asm_nb), ASM_B, 0, ctx->stream>>>(
        ctx->n, ctx->Fi, asm_view(ctx), ctx->cf_ptr.p, ctx->cf_face.p, ctx->owner.p,
        ctx->neighbour.p, geom_of(ctx), mu, mu_f, bc_kind, phi, ctx->slot_pn.p, ctx->slot_np.p,
        ctx->ia.p, ctx->diag.p, ctx->a.p, su);
  else
    k_laplacian<<<xcd_grid(ctx->n), BLK, 0, ctx->stream>>>(
        ctx->n, ctx->Fi, ctx->cf_ptr.p, ctx->cf_face.p, ctx->owner.p, ctx->neighbour.p,
        geom_of(ctx), mu, mu_f, bc_kind, phi, ctx->slot_pn.p, ctx->slot_np.p, ctx->diag.p,
        ctx->a.p, su);
  if (ctx->nper > 0)
    k_zero_periodic<<<grid_for(ctx->nper, BLK), BLK, 0, ctx->stream>>>(
        ctx->Fi, ctx->nper, ctx->slot_pn.p, ctx->slot_np.p, ctx->a.p);
  FCP_HIP(ctx, hipGetLastError());
  ctx->a_version++;
  return fcp_coarse_begin(ctx);
}

int fcp_launch_nonorth(fcp_ctx *ctx, const double *gamma_f, const double *grad, double *su) {
  if (asm_on(ctx)) {
    ProfScope ps(ctx, "nonorth");
    k_nonorth_blk<<<asm_grid(ctx->asm_nb), ASM_B, 0, ctx->stream>>>(
        ctx->n, ctx->Fi, asm_view(ctx), ctx->cf_ptr.p, ctx->cf_face.p, ctx->owner.p,
        ctx->neighbour.p, geom_of(ctx), gamma_f, grad, su);
    FCP_HIP(ctx, hipGetLastError());
    return 0;
  }
  ProfScope ps(ctx, "nonorth");
  k_nonorth<<<xcd_grid(ctx->n), BLK, 0, ctx->stream>>>(ctx->n, ctx->Fi, ctx->cf_ptr.p,
                                                              ctx->cf_face.p, ctx->owner.p,
                                                              ctx->neighbour.p, geom_of(ctx),
                                                              gamma_f, grad, su);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

int fcp_launch_grad_gauss(fcp_ctx *ctx, const double *u, double *dudxi) {
  const int ntot = ctx->n + ctx->Fb;
  if (asm_on(ctx)) {
    ProfScope ps(ctx, "grad_gauss");
    k_grad_blk<<<asm_grid(ctx->asm_nb), ASM_B, 0, ctx->stream>>>(
        ctx->n, ctx->Fi, asm_view(ctx), ctx->cf_ptr.p, ctx->cf_face.p, ctx->owner.p,
        ctx->neighbour.p, geom_of(ctx), u, dudxi);
    if (ctx->Fb > 0)
      k_zero_rows3<<<grid_for(ctx->Fb, BLK), BLK, 0, ctx->stream>>>(ctx->n, ntot, dudxi);
    FCP_HIP(ctx, hipGetLastError());
    return 0;
  }
  ProfScope ps(ctx, "grad_gauss");
  k_grad_gauss<<<xcd_grid(ntot), BLK, 0, ctx->stream>>>(
      ctx->n, ctx->Fi, ntot, ctx->cf_ptr.p, ctx->cf_face.p, ctx->owner.p, ctx->neighbour.p,
      geom_of(ctx), u, dudxi);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

int fcp_launch_grad_fvx(fcp_ctx *ctx, const double *u, const double *dfo, double *tmp,
                        double *dudxi) {
  ProfScope ps(ctx, "grad_fvx");
  const int ntot = ctx->n + ctx->Fb;
  auto sweep = [&](const double *old, double *out) {
    k_grad_fvx<<<xcd_grid(ntot), BLK, 0, ctx->stream>>>(
        ctx->n, ctx->Fi, ntot, ctx->cf_ptr.p, ctx->cf_face.p, ctx->owner.p, ctx->neighbour.p,
        geom_of(ctx), u, old, out);
  };
  if (dfo) {
    sweep(dfo, dudxi);
  } else {
    sweep(nullptr, tmp);
    sweep(tmp, dudxi);
  }
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

int fcp_launch_face_gamma_cells(fcp_ctx *ctx, const double *dc, double *gf) {
  const int F = ctx->F;
  if (F == 0) return 0;
  k_face_gamma_cells<<<grid_for(F, BLK), BLK, 0, ctx->stream>>>(
      ctx->Fi, F, ctx->owner.p, ctx->neighbour.p, ctx->facint.p, dc, gf);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

int fcp_launch_update_boundary(fcp_ctx *ctx, double *phi) {
  if (ctx->Fb == 0) return 0;
  k_update_boundary<<<grid_for(ctx->Fb, BLK), BLK, 0, ctx->stream>>>(
      ctx->n, ctx->Fi, ctx->Fb, ctx->owner.p, ctx->bpatch.p, ctx->bctype.p,
      fcp_periodic_view(ctx), phi);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

int fcp_launch_convdiff(fcp_ctx *ctx, const double *flmass, const double *dcoef, double *phi,
                        const double *grad, double gam, int scheme, const unsigned char *bc_kind,
                        double urf, double *su, double *sp, const int *srd_map, int cn,
                        const double *den, double dt, const double *phi_o) {
  FCP_TRY(fcp_coarse_join(ctx));
  ProfScope ps(ctx, "convdiff");
  k_convdiff<<<xcd_grid(ctx->n), BLK, 0, ctx->stream>>>(
      ctx->n, ctx->Fi, ctx->cf_ptr.p, ctx->cf_face.p, ctx->owner.p, ctx->neighbour.p, geom_of(ctx),
      ctx->srdb.p, flmass, dcoef, phi, grad, gam, scheme, bc_kind, urf, ctx->slot_pn.p,
      ctx->slot_np.p, ctx->ia.p, ctx->diag.p, ctx->a.p, su, sp, srd_map, cn, den, dt, phi_o,
      fcp_periodic_view(ctx));
  FCP_HIP(ctx, hipGetLastError());
  ctx->a_version++;
  return 0;
}

int fcp_launch_ddt(fcp_ctx *ctx, int scheme, const double *den, double dt, const double *fo,
                   const double *foo, double *su, double *sp) {
  k_ddt<<<grid_for(ctx->n, BLK), BLK, 0, ctx->stream>>>(ctx->n, scheme, den, ctx->vol.p, dt, fo, foo,
                                                        su, sp);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

int fcp_launch_transport_dcoef(fcp_ctx *ctx, const double *vis, const unsigned char *bc_kind,
                               double viscos, double pranl, double sigt, double *dcoef) {
  const double prl = 1.0 / pranl, prtr = 1.0 / sigt;  // temperature.f90:66-68
  k_transport_dcoef<<<grid_for(ctx->F, BLK), BLK, 0, ctx->stream>>>(
      ctx->Fi, ctx->F, ctx->owner.p, ctx->neighbour.p, ctx->facint.p, vis, bc_kind, viscos, prl,
      prtr, pranl, sigt, fcp_periodic_view(ctx), dcoef);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

int fcp_launch_minmax_clip(fcp_ctx *ctx, double *x, int clip, double *mm) {
  const int nb = grid_for(ctx->n, BLK) < 1024 ? grid_for(ctx->n, BLK) : 1024;
  k_minmax<<<nb, BLK, 0, ctx->stream>>>(ctx->n, x, mm);
  k_minmax_final<<<1, 64, 0, ctx->stream>>>(nb, mm);
  if (clip) k_clip_small<<<grid_for(ctx->n, BLK), BLK, 0, ctx->stream>>>(ctx->n, x, mm);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

int fcp_launch_pcorr_assemble(fcp_ctx *ctx, double *u, double *v, double *w, const double *p,
                              const double *dpdxi, const double *den, const double *apu,
                              const double *apv, const double *apw, double *flmass,
                              int const_mflux, double flomas, double *su) {
  FCP_TRY(fcp_coarse_join(ctx));
  ProfScope ps(ctx, "pcorr_assemble");
  const int n = ctx->n, Fi = ctx->Fi, Fb = ctx->Fb;
  const GeomP g = geom_of(ctx);
  if (!const_mflux && Fb > 0) {  // adjustMassFlow
    FCP_HIP(ctx, ctx->pc_fac.alloc(1));
    k_pc_outlet<<<grid_for(Fb, BLK), BLK, 0, ctx->stream>>>(n, Fi, Fb, ctx->owner.p, ctx->bpatch.p,
                                                            ctx->bctype.p, g, den, u, v, w, flmass);
    k_pc_flowo<<<1, 256, 0, ctx->stream>>>(Fi, Fb, ctx->bpatch.p, ctx->bctype.p, flmass, flomas,
                                           ctx->pc_fac.p);
    k_pc_scale<<<grid_for(Fb, BLK), BLK, 0, ctx->stream>>>(n, Fi, Fb, ctx->bpatch.p, ctx->bctype.p,
                                                           ctx->pc_fac.p, u, v, w, flmass);
  }
  const RcIn in{u, v, w, p, dpdxi, den, apu, apv, apw};
  k_pc_assemble<<<xcd_grid(n), BLK, 0, ctx->stream>>>(
      n, Fi, ctx->cf_ptr.p, ctx->cf_face.p, ctx->owner.p, ctx->neighbour.p, ctx->bpatch.p,
      ctx->bctype.p, g, in, ctx->slot_pn.p, ctx->slot_np.p, ctx->diag.p, flmass, ctx->a.p, su,
      fcp_periodic_view(ctx));
  FCP_HIP(ctx, hipGetLastError());
  ctx->a_version++;
  return 0;
}

int fcp_launch_pcorr_correct(fcp_ctx *ctx, double *pp, double *flmass, double *u, double *v,
                             double *w, double *p, double *dpdxi, const double *apu,
                             const double *apv, const double *apw, double urfp, int pref,
                             double *su, double *sv, double *sw) {
  ProfScope ps(ctx, "pcorr_correct");
  const int n = ctx->n, Fi = ctx->Fi, Fb = ctx->Fb, F = ctx->F;
  const GeomP g = geom_of(ctx);
  if (F > 0)
    k_pc_fluxcorr<<<grid_for(F, BLK), BLK, 0, ctx->stream>>>(Fi, F, ctx->owner.p,
                                                             ctx->neighbour.p, ctx->slot_pn.p,
                                                             ctx->a.p, pp, flmass,
                                                             fcp_periodic_view(ctx));
  for (int stage = 1; stage <= 2; ++stage) {
    if (Fb > 0)
      k_pc_bpres<<<grid_for(Fb, BLK), BLK, 0, ctx->stream>>>(n, Fi, Fb, stage, ctx->owner.p,
                                                             ctx->bpatch.p, ctx->bctype.p, g,
                                                             dpdxi, pp);
    k_pc_grad<<<xcd_grid(n), BLK, 0, ctx->stream>>>(
        n, Fi, ctx->cf_ptr.p, ctx->cf_face.p, ctx->owner.p, ctx->neighbour.p, g, pp, stage == 1,
        stage == 2, dpdxi, su, sv, sw, apu, apv, apw, urfp, pref, u, v, w, p);
  }
  if (Fb > 0)
    k_pc_velbnd<<<grid_for(Fb, BLK), BLK, 0, ctx->stream>>>(n, Fi, Fb, ctx->owner.p, ctx->bpatch.p,
                                                            ctx->bctype.p, g, u, v, w);
  FCP_HIP(ctx, hipGetLastError());
  return 0;
}

int fcp_launch_momentum(fcp_ctx *ctx, double *u, double *v, double *w, const double *uo,
                        const double *vo, const double *wo, const double *uoo,
                        const double *voo, const double *woo, double *p, const double *den,
                        const double *vis, const double *visw, const int *wall_of,
                        const double *flmass, double viscos, double dt, int ts, int scheme,
                        double gds, const double *urf, double *dpdxi, double *grad9,
                        double *s3, double *rhs3, double *diag3, double *ap3, double *sp3,
                        double *src3) {
  FCP_TRY(fcp_coarse_join(ctx));
  ProfScope ps(ctx, "momentum");
  const int n = ctx->n, Fi = ctx->Fi, Fb = ctx->Fb;
  const size_t nt = (size_t)n + Fb;
  const GeomP g = geom_of(ctx);
  if (Fb > 0)  // updateVelocityAtBoundary (:166)
    k_pc_velbnd<<<grid_for(Fb, BLK), BLK, 0, ctx->stream>>>(n, Fi, Fb, ctx->owner.p, ctx->bpatch.p,
                                                            ctx->bctype.p, g, u, v, w);
  double *U[3] = {u, v, w};
  for (int r = 0; r < 3; ++r) FCP_TRY(fcp_launch_grad_gauss(ctx, U[r], grad9 + 3 * nt * r));
  for (int stage = 1; stage <= 2; ++stage) {  // gradp_and_sources(p) (:173)
    if (Fb > 0)
      k_pc_bpres<<<grid_for(Fb, BLK), BLK, 0, ctx->stream>>>(n, Fi, Fb, stage, ctx->owner.p,
                                                             ctx->bpatch.p, ctx->bctype.p, g,
                                                             dpdxi, p);
    k_pc_grad<<<xcd_grid(n), BLK, 0, ctx->stream>>>(
        n, Fi, ctx->cf_ptr.p, ctx->cf_face.p, ctx->owner.p, ctx->neighbour.p, g, p, stage == 1,
        stage == 2, dpdxi, s3, s3 + n, s3 + 2 * (size_t)n, nullptr, nullptr, nullptr, 0.0, 0,
        nullptr, nullptr, nullptr, nullptr);
  }
  MomIn in;
  for (int r = 0; r < 3; ++r) {
    in.U[r] = U[r];
    in.G[r] = grad9 + 3 * nt * r;
    in.urf[r] = urf[r];
  }
  in.O[0] = uo, in.O[1] = vo, in.O[2] = wo;
  in.OO[0] = uoo, in.OO[1] = voo, in.OO[2] = woo;
  in.den = den, in.vis = vis, in.visw = visw, in.flmass = flmass, in.wall_of = wall_of;
  in.viscos = viscos, in.dt = dt, in.gds = gds, in.ts = ts, in.scheme = scheme;
  k_mom_cells<<<xcd_grid(n), BLK, 0, ctx->stream>>>(
      n, Fi, ctx->cf_ptr.p, ctx->cf_face.p, ctx->owner.p, ctx->neighbour.p, ctx->bpatch.p,
      ctx->bctype.p, g, in, ctx->slot_pn.p, ctx->slot_np.p, ctx->ia.p, ctx->diag.p, ctx->a.p, s3,
      rhs3, diag3, ap3, sp3, src3, fcp_periodic_view(ctx));
  FCP_HIP(ctx, hipGetLastError());
  ctx->a_version++;
  return 0;
}
