// ORACLE — test infrastructure only (see orc_common.h).
// LST::prepareInput (SDL/LST.cc:163-365): seeds -> pLS pseudo-hits,
// superbins and pixel types, in double precision as ROOT::Math::XYZVector
// computes them.  The XYZVector eta() used here is ROOT GenVector's
// Eta_FromRhoZ (rho>0 branch: log(z/rho + sqrt((z/rho)^2 + 1))); ROOT is not
// available in this image, so that one routine is "parity unpinned" against
// ROOT itself and pinned instead by the closed form asinh(z/rho).
#include "orc_prepare.h"
#include <cmath>
#include <limits>

namespace orc {
namespace {
struct V3 {
  double x, y, z;
  double rho() const { return std::sqrt(x * x + y * y); }
  double r() const { return std::sqrt(x * x + y * y + z * z); }
  double phi() const { return (x == 0.0 && y == 0.0) ? 0.0 : std::atan2(y, x); }
  double eta() const {
    double rh = rho();
    if (rh > 0) {
      static const double big = std::pow(std::numeric_limits<double>::epsilon(), -0.25);
      double zs = z / rh;
      if (std::fabs(zs) < big) return std::log(zs + std::sqrt(zs * zs + 1.0));
      return z > 0 ? std::log(2.0 * zs + 0.5 / zs) : -std::log(-2.0 * zs);
    }
    if (z == 0) return 0;
    return z > 0 ? z + 22756.0 : z - 22756.0;
  }
};
// SDL/LST.cc:163-172
V3 r3FromPCA(const V3& p3, float dxy, float dz) {
  const float pt = p3.rho();
  const float p = p3.r();
  const float vz = dz * pt * pt / p / p;
  const float vx = -dxy * p3.y / pt - p3.x / p * p3.z / p * dz;
  const float vy = dxy * p3.x / pt - p3.y / p * p3.z / p * dz;
  return {vx, vy, vz};
}
}  // namespace

void prepare_input(const SeedInput& s, double plsPtCut, Input& out) {
  out = Input();
  const size_t nHit = s.ph2_x.size();
  out.x = s.ph2_x; out.y = s.ph2_y; out.z = s.ph2_z; out.detId = s.ph2_detId;
  out.idx.resize(nHit);
  for (size_t i = 0; i < nHit; i++) out.idx[i] = i;
  const int hit_size = nHit;
  unsigned count = 0;
  const size_t n_see = s.see_stateTrajGlbPx.size();
  for (size_t i = 0; i < n_see; i++) {
    V3 p3LH{s.see_stateTrajGlbPx[i], s.see_stateTrajGlbPy[i], s.see_stateTrajGlbPz[i]};
    float ptIn = p3LH.rho();
    float eta = p3LH.eta();
    float ptErr = s.see_ptErr[i];
    if (!(ptIn > plsPtCut - 2 * ptErr)) continue;
    V3 r3LH{s.see_stateTrajGlbX[i], s.see_stateTrajGlbY[i], s.see_stateTrajGlbZ[i]};
    V3 p3PCA{s.see_px[i], s.see_py[i], s.see_pz[i]};
    V3 r3PCA = r3FromPCA(p3PCA, s.see_dxy[i], s.see_dz[i]);
    V3 diff{r3LH.x - p3LH.x, r3LH.y - p3LH.y, r3LH.z - p3LH.z};
    float dPhiChange = diff.phi();  // SDL/LST.cc:250 (the API's definition, kept as is)
    int pixtype;
    if (ptIn >= 2.0) pixtype = 0;
    else if (ptIn >= (plsPtCut - 2 * ptErr) && ptIn < 2.0) pixtype = dPhiChange >= 0 ? 1 : 2;
    else continue;
    const int nh = s.see_hitIdx_off[i + 1] - s.see_hitIdx_off[i];
    const int* hi = s.see_hitIdx.data() + s.see_hitIdx_off[i];
    unsigned h0 = hit_size + count++;
    unsigned h1 = hit_size + count++;
    unsigned h2 = hit_size + count++;
    unsigned h3;
    if (nh <= 3) h3 = h2;
    else h3 = hit_size + count++;
    out.x.push_back(r3PCA.x); out.y.push_back(r3PCA.y); out.z.push_back(r3PCA.z);
    out.x.push_back(p3PCA.rho());
    float p3PCA_Eta = p3PCA.eta();
    out.y.push_back(p3PCA_Eta);
    float p3PCA_Phi = p3PCA.phi();
    out.z.push_back(p3PCA_Phi);
    out.x.push_back(r3LH.x); out.y.push_back(r3LH.y); out.z.push_back(r3LH.z);
    out.detId.push_back(1); out.detId.push_back(1); out.detId.push_back(1);
    if (nh > 3) {
      out.x.push_back(r3LH.x); out.y.push_back(s.see_dxy[i]); out.z.push_back(s.see_dz[i]);
      out.detId.push_back(1);
    }
    out.px.push_back(p3LH.x); out.py.push_back(p3LH.y); out.pz.push_back(p3LH.z);
    out.hitIdx0.push_back(h0); out.hitIdx1.push_back(h1); out.hitIdx2.push_back(h2); out.hitIdx3.push_back(h3);
    out.ptIn.push_back(ptIn); out.ptErr.push_back(ptErr); out.etaErr.push_back(s.see_etaErr[i]);
    out.eta.push_back(eta);
    float phi = p3LH.phi();
    out.phi.push_back(phi);
    out.charge.push_back(s.see_q[i]);
    out.seedIdx.push_back(i);
    out.dPhiChange.push_back(dPhiChange);
    out.idx.push_back(hi[0]); out.idx.push_back(hi[1]); out.idx.push_back(hi[2]);
    char isQuad = 0;
    if (nh > 3) { isQuad = 1; out.idx.push_back(hi[3]); }
    // superbin: SDL/LST.cc:326-335
    float neta = 25., nphi = 72., nz = 25.;
    int etabin = (p3PCA_Eta + 2.6) / ((2 * 2.6) / neta);
    int phibin = (p3PCA_Phi + 3.14159265358979323846) / ((2. * 3.14159265358979323846) / nphi);
    int dzbin = (s.see_dz[i] + 30) / (2 * 30 / nz);
    int isuperbin = (nz * nphi) * etabin + (nz)*phibin + dzbin;
    out.superbin.push_back(isuperbin);
    out.pixelType.push_back(pixtype);
    out.isQuad.push_back(isQuad);
  }
}

}  // namespace orc
