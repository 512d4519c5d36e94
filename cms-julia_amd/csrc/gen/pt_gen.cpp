// pt_gen.cpp - seeded synthetic events for the pixel-track stage (see
// include/pixeltrack_gen.h).  Deterministic for a given seed on a given libm.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../../include/pixeltrack.h"
#include "../../../include/pixeltrack_gen.h"

namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kBarrelR[4] = {2.9, 6.8, 10.9, 16.0};
constexpr double kBarrelHalfZ = 26.6;
constexpr double kDiskZ[3] = {29.1, 39.6, 51.6};
constexpr double kDiskRmin = 4.5, kDiskRmax = 16.0, kDiskRring = 10.25;
constexpr int kLadders[4] = {12, 28, 44, 64};
constexpr int kModuleStart[11] = {0, 96, 320, 672, 1184, 1296, 1408, 1520, 1632, 1744, 1856};
constexpr int kSectors = 56;

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0xD1B54A32D192ED03ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        double u1 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
};

// approx_atan2.jl:164-184 (degree-7 polynomial coefficients as used there)
int16_t unsafe_atan2s(float y, float x) {
    float r = (std::fabs(x) - std::fabs(y)) / (std::fabs(x) + std::fabs(y));
    if (x < 0) r = -r;
    float z = r * r;
    float p = r * (-10422.177734375f + z * (3349.97412109375f + z * (-1525.589599609375f + z * 406.64190673828125f)));
    int16_t angle = (x >= 0) ? (int16_t)8192 : (int16_t)24576;
    angle = (int16_t)(angle + (int16_t)std::trunc(p));
    return (y < 0) ? (int16_t)(-angle) : angle;
}

struct ModuleFrame {
    double ex[3], ey[3], ez[3];
};

ModuleFrame frame_of(int det) {
    ModuleFrame f{};
    int layer = 0;
    while (layer < 10 && det >= kModuleStart[layer + 1]) ++layer;
    const int local = det - kModuleStart[layer];
    double phic;
    if (layer < 4) {
        const int ladder = local / 8;
        phic = -kPi + (ladder + 0.5) * 2.0 * kPi / kLadders[layer];
        const double s = std::sin(phic), c = std::cos(phic);
        f = {{-s, c, 0.0}, {0.0, 0.0, 1.0}, {c, s, 0.0}};
    } else {
        const int sector = local % kSectors;
        phic = -kPi + (sector + 0.5) * 2.0 * kPi / kSectors;
        const double s = std::sin(phic), c = std::cos(phic);
        f = {{-s, c, 0.0}, {c, s, 0.0}, {0.0, 0.0, -1.0}};
    }
    return f;
}

int barrel_module(int layer, double phi, double z) {
    int lad = (int)std::floor((phi + kPi) / (2.0 * kPi) * kLadders[layer]);
    lad = ((lad % kLadders[layer]) + kLadders[layer]) % kLadders[layer];
    int zm = (int)std::floor((z + kBarrelHalfZ) / (2.0 * kBarrelHalfZ / 8.0));
    zm = std::min(7, std::max(0, zm));
    return kModuleStart[layer] + lad * 8 + zm;
}
int disk_module(int layer, double phi, double r) {
    int sec = (int)std::floor((phi + kPi) / (2.0 * kPi) * kSectors);
    sec = ((sec % kSectors) + kSectors) % kSectors;
    const int ring = (r >= kDiskRring) ? 1 : 0;
    return kModuleStart[layer] + ring * kSectors + sec;
}

struct Helix {
    double pt, phi0, eta, z0, R, cot;
    int q;
    // position at transverse arc length s
    void at(double s, double &x, double &y, double &z) const {
        const double cx = q * R * std::sin(phi0), cy = -q * R * std::cos(phi0);
        const double rx = -cx, ry = -cy;
        const double a = -q * s / R;
        const double ca = std::cos(a), sa = std::sin(a);
        x = cx + ca * rx - sa * ry;
        y = cy + sa * rx + ca * ry;
        z = z0 + s * cot;
    }
};

struct RawHit {
    int layer;
    int det;
    double x, y, z;
    double s;  // path order
    int16_t ysize;
    float var_x, var_y;
    int32_t track;
};

// Crossings of a prompt helix with the layers, ordered by path length.
void crossings(const Helix &h, std::vector<RawHit> &out, int track) {
    out.clear();
    const double smax = kPi * h.R;  // first half turn only
    for (int L = 0; L < 4; ++L) {
        const double r = kBarrelR[L];
        if (r >= 2.0 * h.R) continue;
        const double s = 2.0 * h.R * std::asin(r / (2.0 * h.R));
        double x, y, z;
        h.at(s, x, y, z);
        if (std::fabs(z) > kBarrelHalfZ) continue;
        RawHit rh{};
        rh.layer = L;
        rh.x = x, rh.y = y, rh.z = z, rh.s = s;
        rh.track = track;
        out.push_back(rh);
    }
    if (h.cot != 0.0) {
        for (int d = 0; d < 3; ++d) {
            const double zd = (h.cot > 0) ? kDiskZ[d] : -kDiskZ[d];
            const double s = (zd - h.z0) / h.cot;
            if (s <= 0.0 || s >= smax) continue;
            double x, y, z;
            h.at(s, x, y, z);
            const double r = std::sqrt(x * x + y * y);
            if (r < kDiskRmin || r > kDiskRmax) continue;
            RawHit rh{};
            rh.layer = (h.cot > 0) ? 4 + d : 7 + d;
            rh.x = x, rh.y = y, rh.z = zd, rh.s = s;
            rh.track = track;
            out.push_back(rh);
        }
    }
    std::sort(out.begin(), out.end(), [](const RawHit &a, const RawHit &b) { return a.s < b.s; });
}

// Assign module, smear in the module frame, fill local variances and cluster size.
void finish_hit(RawHit &rh, Rng &rng, double srphi, double sz, double cot_local) {
    const double phi = std::atan2(rh.y, rh.x);
    if (rh.layer < 4)
        rh.det = barrel_module(rh.layer, phi, rh.z);
    else
        rh.det = disk_module(rh.layer, phi, std::sqrt(rh.x * rh.x + rh.y * rh.y));
    const ModuleFrame f = frame_of(rh.det);
    const double g1 = rng.normal(), g2 = rng.normal();
    for (int c = 0; c < 3; ++c) {
        const double d = srphi * g1 * f.ex[c] + sz * g2 * f.ey[c];
        if (c == 0) rh.x += d;
        if (c == 1) rh.y += d;
        if (c == 2) rh.z += d;
    }
    rh.var_x = (float)(srphi * srphi);
    rh.var_y = (float)(sz * sz);
    // cluster length along local y in 1/8-pixel units: 8 (one pixel) plus the
    // track's projected length, 15.2 per unit |dz/dr| on the barrel (the
    // dz_dr_fact of gpu_pixel_doublets.jl:134) and per unit |dr/dz| on disks
    const double proj = (rh.layer < 4) ? std::fabs(cot_local) : (cot_local != 0.0 ? 1.0 / std::fabs(cot_local) : 50.0);
    double ys = 8.0 + 15.2 * proj * (1.0 + 0.05 * rng.normal());
    ys = std::min(1023.0, std::max(1.0, std::floor(ys + 0.5)));
    rh.ysize = (int16_t)ys;
}

RawHit noise_hit(int layer, Rng &rng) {
    RawHit rh{};
    rh.layer = layer;
    rh.track = -1;
    const double phi = rng.uniform(-kPi, kPi);
    if (layer < 4) {
        const double r = kBarrelR[layer];
        rh.x = r * std::cos(phi);
        rh.y = r * std::sin(phi);
        rh.z = rng.uniform(-kBarrelHalfZ, kBarrelHalfZ);
    } else {
        const double r = std::sqrt(rng.uniform(kDiskRmin * kDiskRmin, kDiskRmax * kDiskRmax));
        rh.x = r * std::cos(phi);
        rh.y = r * std::sin(phi);
        rh.z = (layer < 7) ? kDiskZ[layer - 4] : -kDiskZ[layer - 7];
    }
    return rh;
}

double sample_pt(const ptg_config &c, Rng &rng) {
    if (c.pt_mode == PTG_PT_POWERLAW) {
        const double u = 1.0 - rng.uniform();
        double pt = c.pt_min * std::pow(u, -1.0 / (c.pt_index - 1.0));
        if (c.pt_max > 0) pt = std::min(pt, c.pt_max);
        return pt;
    }
    return std::exp(rng.uniform(std::log(c.pt_min), std::log(c.pt_max)));
}

inline int hist_bin(int16_t t) { return (int)(((uint16_t)t) >> 9) + 1; }

// ---------------------------------------------------------------- multiple scattering
// A track between two deflections: position, direction (azimuth, cot(theta))
// and transverse radius of curvature; q = +1 turns clockwise seen from +z.
struct Flight {
    double x, y, z, phi, cot, R;
    int q;
    void centre(double &cx, double &cy) const {
        cx = x + q * R * std::sin(phi);
        cy = y - q * R * std::cos(phi);
    }
};

// Smallest positive transverse arc length from the current point to layer L
// (barrel cylinder or disk plane) along the current helix; -1 if not reached.
double arc_to_layer(const Flight &f, int L) {
    if (L >= 4) {
        const double zd = (L < 7) ? kDiskZ[L - 4] : -kDiskZ[L - 7];
        if (f.cot == 0.0) return -1.0;
        const double s = (zd - f.z) / f.cot;
        return s > 1e-9 ? s : -1.0;
    }
    // |c + Rot(a)(p - c)|^2 = r^2 with a = -q s / R: cos(a + delta) = K
    const double r = kBarrelR[L];
    double cx, cy;
    f.centre(cx, cy);
    const double c2 = cx * cx + cy * cy, cn = std::sqrt(c2);
    if (cn == 0.0) return -1.0;
    const double K = (r * r - c2 - f.R * f.R) / (2.0 * cn * f.R);
    if (K < -1.0 || K > 1.0) return -1.0;
    const double delta = std::atan2(f.y - cy, f.x - cx) - std::atan2(cy, cx);
    const double ac = std::acos(K);
    double best = -1.0;
    for (int sg = -1; sg <= 1; sg += 2)
        for (int m = -3; m <= 3; ++m) {
            const double s = -f.q * f.R * (sg * ac - delta + 2.0 * kPi * m);
            if (s > 1e-9 && (best < 0.0 || s < best)) best = s;
        }
    return best;
}

void advance(Flight &f, double s) {
    double cx, cy;
    f.centre(cx, cy);
    const double a = -f.q * s / f.R;
    const double ux = f.x - cx, uy = f.y - cy, ca = std::cos(a), sa = std::sin(a);
    f.x = cx + ca * ux - sa * uy;
    f.y = cy + sa * ux + ca * uy;
    f.z += s * f.cot;
    f.phi += a;
}

// Highland's width of the projected scattering angle (PDG 2022, eq. 34.15;
// beta = 1, unit charge) squared, for a path of `len` cm through the medium
// the broken-line fit assumes: 1/X0 = 0.06/16 per cm and a geometry factor
// 0.7 on the variance (broken_line.jl:46-79).
double highland_var(double len, double p) {
    const double x = len * (0.06 / 16.0);
    const double f = 1.0 + 0.038 * std::log(x);
    return 0.7 * (0.0136 / p) * (0.0136 / p) * x * f * f;
}

// Deflects the direction by Gaussian angles of variance v_bend (towards the
// transverse normal of the direction) and v_polar (in the plane of the
// direction and the z axis); |p| is kept.
void kink(Flight &f, double p, double kappa, double v_bend, double v_polar, Rng &rng) {
    const double st = 1.0 / std::sqrt(1.0 + f.cot * f.cot), ct = f.cot * st;
    const double cp = std::cos(f.phi), sp = std::sin(f.phi);
    const double d1 = v_bend > 0.0 ? std::sqrt(v_bend) * rng.normal() : 0.0;
    const double d2 = v_polar > 0.0 ? std::sqrt(v_polar) * rng.normal() : 0.0;
    const double tx = cp * st - d1 * sp + d2 * ct * cp;
    const double ty = sp * st + d1 * cp + d2 * ct * sp;
    const double tz = ct - d2 * st;
    const double tn = std::sqrt(tx * tx + ty * ty + tz * tz), tr = std::sqrt(tx * tx + ty * ty);
    f.phi = std::atan2(ty, tx);
    f.cot = tz / tr;
    f.R = p * (tr / tn) / kappa;
}

// Positions of the first n crossings of `layers` for a helix deflected as
// described at ptg_fit_config.ms_scale; false if a layer is missed.
bool scattered_hits(const Helix &h, double kappa, double ms_scale, const std::vector<RawHit> &layers, int n,
                    Rng &rng, std::vector<RawHit> &out) {
    const double p = h.pt * std::sqrt(1.0 + h.cot * h.cot);
    Flight f{0.0, 0.0, h.z0, h.phi0, h.cot, h.R, h.q};
    auto path = [](double s, const Flight &g) { return s * std::sqrt(1.0 + g.cot * g.cot); };
    out.assign(layers.begin(), layers.begin() + n);
    for (int i = 0; i < n; ++i) {
        const double s = arc_to_layer(f, layers[i].layer);
        if (s < 0.0) return false;
        const double prev_len = path(s, f);
        advance(f, s);
        out[i].x = f.x, out[i].y = f.y, out[i].z = f.z;
        if (i + 1 < n) {
            // the material of each segment between hits acts at both of its
            // ends: the first hit takes the next segment's share, an interior
            // hit both (the next segment along the current direction)
            Flight g = f;
            const double sn = arc_to_layer(g, layers[i + 1].layer);
            if (sn < 0.0) return false;
            double v = highland_var(path(sn, g), p);
            if (i > 0) v += highland_var(prev_len, p);
            v *= ms_scale;
            kink(f, p, kappa, v, v, rng);
        }
    }
    return true;
}

template <class T> T *dup_array(const std::vector<T> &v) {
    T *p = (T *)std::malloc(std::max<size_t>(1, v.size()) * sizeof(T));
    if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

}  // namespace

extern "C" {

void ptg_config_default(ptg_config *c) {
    c->n_helices = 100;
    c->n_vertices = 1;
    c->noise_frac = 0.0;
    c->pt_mode = PTG_PT_LOGUNIFORM;
    c->pt_min = 0.5;
    c->pt_max = 10.0;
    c->pt_index = 2.5;
    c->eta_max = 2.5;
    c->z_sigma = 5.0;
    c->b_tesla = 3.8;
    c->sigma_rphi = 10e-4;
    c->sigma_z = 20e-4;
    c->seed = 1;
}

int ptg_frames(float *rot, uint32_t n_modules) {
    if (!rot || n_modules < PT_N_MODULES) return PT_ERR_ARG;
    for (uint32_t m = 0; m < PT_N_MODULES; ++m) {
        const ModuleFrame f = frame_of((int)m);
        for (int c = 0; c < 3; ++c) {
            rot[9 * m + c] = (float)f.ex[c];
            rot[9 * m + 3 + c] = (float)f.ey[c];
            rot[9 * m + 6 + c] = (float)f.ez[c];
        }
    }
    return PT_OK;
}

int ptg_iphi(uint32_t n, const float *x, const float *y, int16_t *iphi) {
    if (n && (!x || !y || !iphi)) return PT_ERR_ARG;
    for (uint32_t i = 0; i < n; ++i) iphi[i] = unsafe_atan2s(y[i], x[i]);
    return PT_OK;
}

int ptg_build_phi_hist(const int16_t *iphi, const uint32_t *layer_start, uint32_t *off, uint16_t *bins) {
    const uint32_t nh = 10;
    for (uint32_t b = 0; b < PT_HIST_NOFF; ++b) off[b] = 0;
    auto hist_of = [&](uint32_t i) -> int {
        uint32_t k = 1;  // searchsortedfirst over the 11 offsets
        while (k <= nh + 1 && layer_start[k - 1] < i) ++k;
        if (k <= nh + 1 && layer_start[k - 1] == i) k += 1;
        return (int)k - 2;
    };
    const uint32_t n = layer_start[nh];
    for (uint32_t i = 0; i < n; ++i) {
        const int ih = hist_of(i);
        if (ih < 0 || ih >= (int)nh) return PT_ERR_ARG;
        off[hist_bin(iphi[i]) + 128 * ih - 1] += 1;
    }
    for (uint32_t b = 1; b < PT_HIST_NOFF; ++b) off[b] += off[b - 1];
    for (uint32_t i = 0; i < n; ++i) {
        const int ih = hist_of(i);
        const uint32_t b = hist_bin(iphi[i]) + 128 * ih;
        const uint32_t w = off[b - 1];
        off[b - 1] -= 1;
        bins[w - 1] = (uint16_t)(i + 1);
    }
    return PT_OK;
}

int ptg_generate_event(const ptg_config *c, ptg_event **out) {
    if (!c || !out) return PT_ERR_ARG;
    Rng rng(c->seed);
    const double kappa = 0.299792458e-2 * c->b_tesla;  // GeV/cm
    std::vector<double> vz(std::max<uint32_t>(1, c->n_vertices));
    for (auto &z : vz) z = c->z_sigma * rng.normal();
    std::vector<RawHit> hits, cross;
    hits.reserve((size_t)c->n_helices * 6);
    ptg_event *e = (ptg_event *)std::calloc(1, sizeof(ptg_event));
    std::vector<float> tpt, tphi, teta, tz0;
    std::vector<int8_t> tq;
    for (uint32_t t = 0; t < c->n_helices; ++t) {
        Helix h;
        h.pt = sample_pt(*c, rng);
        h.phi0 = rng.uniform(-kPi, kPi);
        h.eta = rng.uniform(-c->eta_max, c->eta_max);
        h.q = (rng.uniform() < 0.5) ? -1 : 1;
        h.z0 = vz[(size_t)(rng.uniform() * vz.size()) % vz.size()];
        h.R = h.pt / kappa;
        h.cot = std::sinh(h.eta);
        tpt.push_back((float)h.pt);
        tphi.push_back((float)h.phi0);
        teta.push_back((float)h.eta);
        tz0.push_back((float)h.z0);
        tq.push_back((int8_t)h.q);
        crossings(h, cross, (int)t);
        for (auto &rh : cross) {
            finish_hit(rh, rng, c->sigma_rphi, c->sigma_z, h.cot);
            hits.push_back(rh);
        }
    }
    // uniform noise, layer chosen in proportion to the track hits of the layer
    const size_t n_track_hits = hits.size();
    const size_t n_noise = (size_t)std::llround(c->noise_frac * (double)n_track_hits);
    std::vector<double> cum(10, 0.0);
    for (auto &h : hits) cum[h.layer] += 1.0;
    for (int L = 1; L < 10; ++L) cum[L] += cum[L - 1];
    for (size_t k = 0; k < n_noise; ++k) {
        const double u = rng.uniform() * cum[9];
        int L = 0;
        while (L < 9 && u >= cum[L]) ++L;
        RawHit rh = noise_hit(L, rng);
        finish_hit(rh, rng, c->sigma_rphi, c->sigma_z, rng.uniform(-6.0, 6.0));
        hits.push_back(rh);
    }
    // no layer may be empty (an empty layer misfiles the next hit in the histogram)
    uint32_t fillers = 0;
    for (int L = 0; L < 10; ++L) {
        bool has = false;
        for (auto &h : hits) has |= (h.layer == L);
        if (!has) {
            RawHit rh = noise_hit(L, rng);
            finish_hit(rh, rng, c->sigma_rphi, c->sigma_z, 1.0);
            hits.push_back(rh);
            ++fillers;
        }
    }
    std::stable_sort(hits.begin(), hits.end(), [](const RawHit &a, const RawHit &b) { return a.det < b.det; });
    const uint32_t n = (uint32_t)hits.size();
    std::vector<float> xg(n), yg(n), zg(n), rg(n), xe(n), ye(n);
    std::vector<int16_t> ip(n), ys(n), dt(n);
    std::vector<int32_t> tr(n);
    for (uint32_t i = 0; i < n; ++i) {
        const RawHit &h = hits[i];
        xg[i] = (float)h.x;
        yg[i] = (float)h.y;
        zg[i] = (float)h.z;
        rg[i] = std::sqrt(xg[i] * xg[i] + yg[i] * yg[i]);
        xe[i] = h.var_x;
        ye[i] = h.var_y;
        ip[i] = unsafe_atan2s(yg[i], xg[i]);
        ys[i] = h.ysize;
        dt[i] = (int16_t)h.det;
        tr[i] = h.track;
    }
    e->n_hits = n;
    for (int L = 0; L <= 10; ++L) {
        uint32_t first = n;
        for (uint32_t i = 0; i < n; ++i)
            if (dt[i] >= kModuleStart[L]) {
                first = i;
                break;
            }
        e->layer_start[L] = (L == 10) ? n : first;
    }
    e->xg = dup_array(xg);
    e->yg = dup_array(yg);
    e->zg = dup_array(zg);
    e->rg = dup_array(rg);
    e->xerr = dup_array(xe);
    e->yerr = dup_array(ye);
    e->iphi = dup_array(ip);
    e->ysize = dup_array(ys);
    e->det_ind = dup_array(dt);
    e->truth_track = dup_array(tr);
    e->hist_bins = (uint16_t *)std::calloc(std::max<uint32_t>(1, n), sizeof(uint16_t));
    if (ptg_build_phi_hist(e->iphi, e->layer_start, e->hist_off, e->hist_bins) != PT_OK) {
        ptg_free_event(e);
        return PT_ERR_ARG;
    }
    e->n_tracks = c->n_helices;
    e->trk_pt = dup_array(tpt);
    e->trk_phi = dup_array(tphi);
    e->trk_eta = dup_array(teta);
    e->trk_z0 = dup_array(tz0);
    e->trk_q = dup_array(tq);
    e->n_filler = fillers;
    *out = e;
    return PT_OK;
}

void ptg_free_event(ptg_event *e) {
    if (!e) return;
    std::free(e->xg);
    std::free(e->yg);
    std::free(e->zg);
    std::free(e->rg);
    std::free(e->xerr);
    std::free(e->yerr);
    std::free(e->iphi);
    std::free(e->ysize);
    std::free(e->det_ind);
    std::free(e->hist_bins);
    std::free(e->truth_track);
    std::free(e->trk_pt);
    std::free(e->trk_phi);
    std::free(e->trk_eta);
    std::free(e->trk_z0);
    std::free(e->trk_q);
    std::free(e);
}

int ptg_generate_fit_tracks(const ptg_fit_config *c, float *hits, float *ge, uint8_t *nhit, float *sig, int16_t *det,
                            float *truth5) {
    if (!c || !hits || !ge || !nhit) return PT_ERR_ARG;
    Rng rng(c->seed);
    const double kappa = 0.299792458e-2 * c->b_tesla;
    const uint32_t n = c->n_tracks;
    std::vector<RawHit> cross;
    const double m0 = c->len_mix[0], m1 = m0 + c->len_mix[1];
    for (uint32_t t = 0; t < n; ++t) {
        const double u = rng.uniform();
        const int want = (u < m0) ? 3 : (u < m1 ? 4 : 5);
        Helix h;
        for (;;) {
            h.pt = std::exp(rng.uniform(std::log(c->pt_min), std::log(c->pt_max)));
            h.phi0 = rng.uniform(-kPi, kPi);
            h.eta = rng.uniform(-c->eta_max, c->eta_max);
            h.q = (rng.uniform() < 0.5) ? -1 : 1;
            h.z0 = c->z_sigma * rng.normal();
            h.R = h.pt / kappa;
            h.cot = std::sinh(h.eta);
            crossings(h, cross, (int)t);
            if ((int)cross.size() < want) continue;
            if (c->ms_scale > 0.0) {
                std::vector<RawHit> sc;
                if (!scattered_hits(h, kappa, c->ms_scale, cross, want, rng, sc)) continue;
                std::copy(sc.begin(), sc.end(), cross.begin());
            }
            break;
        }
        nhit[t] = (uint8_t)want;
        for (int k = 0; k < 5; ++k) {
            const int kk = std::min(k, want - 1);
            RawHit rh = cross[kk];
            const double f = kk / 4.0;
            const double srphi = c->srphi0 + (c->srphi1 - c->srphi0) * f;
            const double sz = c->sz0 + (c->sz1 - c->sz0) * f;
            if (k < want) finish_hit(rh, rng, srphi, sz, h.cot);
            else rh = RawHit{}, rh.det = 0;
            float g[6];
            const ModuleFrame fr = frame_of(rh.det);
            float rot[9];
            for (int q = 0; q < 3; ++q) {
                rot[q] = (float)fr.ex[q];
                rot[3 + q] = (float)fr.ey[q];
                rot[6 + q] = (float)fr.ez[q];
            }
            {
                const float cxx = rh.var_x, cyy = rh.var_y, cxy = 0.0f;
                const float xx = rot[0], xy = rot[1], xz = rot[2], yx = rot[3], yy = rot[4], yz = rot[5];
                g[0] = xx * (xx * cxx + yx * cxy) + yx * (xx * cxy + yx * cyy);
                g[1] = xx * (xy * cxx + yy * cxy) + yx * (xy * cxy + yy * cyy);
                g[2] = xy * (xy * cxx + yy * cxy) + yy * (xy * cxy + yy * cyy);
                g[3] = xx * (xz * cxx + yz * cxy) + yx * (xz * cxy + yz * cyy);
                g[4] = xy * (xz * cxx + yz * cxy) + yy * (xz * cxy + yz * cyy);
                g[5] = xz * (xz * cxx + yz * cxy) + yz * (xz * cxy + yz * cyy);
            }
            hits[(size_t)(0 * 5 + k) * n + t] = (float)rh.x;
            hits[(size_t)(1 * 5 + k) * n + t] = (float)rh.y;
            hits[(size_t)(2 * 5 + k) * n + t] = (float)rh.z;
            for (int q = 0; q < 6; ++q) ge[(size_t)(q * 5 + k) * n + t] = g[q];
            if (sig) {
                sig[(size_t)(0 * 5 + k) * n + t] = (float)std::sqrt((double)rh.var_x);
                sig[(size_t)(1 * 5 + k) * n + t] = (float)std::sqrt((double)rh.var_y);
            }
            if (det) det[(size_t)k * n + t] = (int16_t)rh.det;
        }
        if (truth5) {
            // (phi, tip, q/pt, cot(theta), zip) of the generating helix
            truth5[(size_t)0 * n + t] = (float)h.phi0;
            truth5[(size_t)1 * n + t] = 0.0f;
            truth5[(size_t)2 * n + t] = (float)(h.q / h.pt);
            truth5[(size_t)3 * n + t] = (float)h.cot;
            truth5[(size_t)4 * n + t] = (float)h.z0;
        }
    }
    return PT_OK;
}

}  // extern "C"
