// Host-side input preparation of SDL::LST::run (SDL/LST.cc:175-365).
//
// The reference computes these quantities with ROOT::Math::XYZVector, i.e. in
// double precision from float inputs; this file does the same arithmetic with
// plain doubles (rho = sqrt(x^2+y^2), r = sqrt(x^2+y^2+z^2), phi = atan2(y, x)
// with phi(0,0) = 0, eta = log(z/rho + sqrt((z/rho)^2 + 1)) for rho > 0 as in
// ROOT GenVector's Eta_FromRhoZ).
#include "prepare.h"

#include <cmath>
#include <limits>

namespace lst {
namespace {

inline double rho(double x, double y) { return std::sqrt(x * x + y * y); }
inline double phi0(double x, double y) { return (x == 0.0 && y == 0.0) ? 0.0 : std::atan2(y, x); }
inline double eta_of(double x, double y, double z) {
  const double r = rho(x, y);
  if (r > 0) {
    static const double big = std::pow(std::numeric_limits<double>::epsilon(), -0.25);
    const double t = z / r;
    if (std::fabs(t) < big) return std::log(t + std::sqrt(t * t + 1.0));
    return z > 0 ? std::log(2.0 * t + 0.5 / t) : -std::log(-2.0 * t);
  }
  if (z == 0) return 0;
  return z > 0 ? z + 22756.0 : z - 22756.0;
}

}  // namespace

lst_pixel_segments PreparedEvent::view() const {
  lst_pixel_segments v;
  v.hit_indices0 = h0.data(); v.hit_indices1 = h1.data(); v.hit_indices2 = h2.data(); v.hit_indices3 = h3.data();
  v.delta_phi_change = dPhiChange.data(); v.pt_in = ptIn.data(); v.pt_err = ptErr.data();
  v.px = px.data(); v.py = py.data(); v.pz = pz.data(); v.eta = eta.data(); v.eta_err = etaErr.data();
  v.phi = phi.data(); v.charge = charge.data(); v.seed_idx = seedIdx.data(); v.superbin = superbin.data();
  v.pixel_type = pixelType.data(); v.is_quad = isQuad.data();
  v.n_pls = (uint32_t)ptIn.size();
  return v;
}

void prepare_event(const lst_input& in, PreparedEvent& o) {
  o = PreparedEvent();
  const uint32_t nh = in.n_hits;
  o.x.assign(in.ph2_x, in.ph2_x + nh);
  o.y.assign(in.ph2_y, in.ph2_y + nh);
  o.z.assign(in.ph2_z, in.ph2_z + nh);
  o.detId.assign(in.ph2_detId, in.ph2_detId + nh);
  o.idx.resize(nh);
  for (uint32_t i = 0; i < nh; i++) o.idx[i] = i;
  const int hitSize = (int)nh;
  unsigned next = 0;
  for (uint32_t s = 0; s < in.n_seeds; s++) {
    const double lx = in.see_stateTrajGlbPx[s], ly = in.see_stateTrajGlbPy[s], lz = in.see_stateTrajGlbPz[s];
    const float ptIn = rho(lx, ly);
    const float eta = eta_of(lx, ly, lz);
    const float ptErr = in.see_ptErr[s];
    if (!(ptIn > 0.8 - 2 * ptErr)) continue;
    const double rx = in.see_stateTrajGlbX[s], ry = in.see_stateTrajGlbY[s], rz = in.see_stateTrajGlbZ[s];
    const double qx = in.see_px[s], qy = in.see_py[s], qz = in.see_pz[s];
    const float dxy = in.see_dxy[s], dz = in.see_dz[s];
    // point of closest approach from (p, dxy, dz): SDL/LST.cc:164-172
    const float pt = rho(qx, qy);
    const float p = std::sqrt(qx * qx + qy * qy + qz * qz);
    const float vz = dz * pt * pt / p / p;
    const float vx = -dxy * qy / pt - qx / p * qz / p * dz;
    const float vy = dxy * qx / pt - qy / p * qz / p * dz;
    const float dPhiChange = phi0(rx - lx, ry - ly);  // SDL/LST.cc:250, as the API defines it
    int pixtype;
    if (ptIn >= 2.0) pixtype = 0;
    else if (ptIn >= (0.8 - 2 * ptErr) && ptIn < 2.0) pixtype = dPhiChange >= 0 ? 1 : 2;
    else continue;
    const uint32_t off = in.see_hitIdx_offsets[s];
    const uint32_t nSeedHits = in.see_hitIdx_offsets[s + 1] - off;
    const bool quad = nSeedHits > 3;
    const uint32_t i0 = hitSize + next++, i1 = hitSize + next++, i2 = hitSize + next++;
    const uint32_t i3 = quad ? hitSize + next++ : i2;
    const float pcaEta = eta_of(qx, qy, qz);
    const float pcaPhi = phi0(qx, qy);
    o.x.push_back(vx); o.y.push_back(vy); o.z.push_back(vz);
    o.x.push_back(rho(qx, qy)); o.y.push_back(pcaEta); o.z.push_back(pcaPhi);
    o.x.push_back(rx); o.y.push_back(ry); o.z.push_back(rz);
    o.detId.insert(o.detId.end(), {1u, 1u, 1u});
    if (quad) {
      o.x.push_back(rx); o.y.push_back(dxy); o.z.push_back(dz);
      o.detId.push_back(1u);
    }
    o.px.push_back(lx); o.py.push_back(ly); o.pz.push_back(lz);
    o.h0.push_back(i0); o.h1.push_back(i1); o.h2.push_back(i2); o.h3.push_back(i3);
    o.ptIn.push_back(ptIn); o.ptErr.push_back(ptErr); o.etaErr.push_back(in.see_etaErr[s]); o.eta.push_back(eta);
    o.phi.push_back((float)phi0(lx, ly));
    o.charge.push_back(in.see_q[s]);
    o.seedIdx.push_back(s);
    o.dPhiChange.push_back(dPhiChange);
    for (uint32_t k = 0; k < (quad ? 4u : 3u); k++) o.idx.push_back((uint32_t)in.see_hitIdx[off + k]);
    // 25 eta x 72 phi x 25 dz superbins: SDL/LST.cc:326-335
    const float nEta = 25., nPhi = 72., nZ = 25.;
    const int etaBin = (pcaEta + 2.6) / ((2 * 2.6) / nEta);
    const int phiBin = (pcaPhi + 3.14159265358979323846) / ((2. * 3.14159265358979323846) / nPhi);
    const int dzBin = (dz + 30) / (2 * 30 / nZ);
    o.superbin.push_back((nZ * nPhi) * etaBin + (nZ)*phiBin + dzBin);
    o.pixelType.push_back((int8_t)pixtype);
    o.isQuad.push_back(quad ? 1 : 0);
  }
}

}  // namespace lst
