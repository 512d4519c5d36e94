// oracle_ca.cpp - TEST INFRASTRUCTURE (the oracle): serial CPU restatement of
// the reference pixel-track stage, used only by tests/, the smoke check and the
// CPU-baseline leg of bench.py.  Never part of the product path.
//
// Restates, in the reference's serial order and float semantics (build with
// -ffp-contract=off, no fast-math):
//   ca_hit_ntuplet_generator.jl:65-97          make_tuples
//   ca_hit_ntuplet_generator_kernels.jl:121-231 reset/build_doublets/launch_kernels/classify
//   gpu_pixel_doublets.jl:19-360               layer pairs, cuts, doublets_from_histo
//   gpu_ca_cell.jl:28-270                      GPUCACell, are_aligned, dca_cut,
//                                              add_outer_neighbor, add_track, find_ntuplets
//   circle_eq.jl:18-93                         CircleEq, dca0
//   gpu_fishbone.jl:4-62                       fish_bone
//   ca_hit_ntuplet_generator_kernels_impl.jl:25-362 kernels
//   hist_to_container.jl:81-270                bin, bulk_fill, bulk_finalize_fill,
//                                              count/fill_direct, finalize!, fill_many_from_vector
//   simple_matrix.jl:1-58                      PreAllocMatrix (extend!/push!/reset!)
//   broken_line_fit_on_gpu.jl:24-203           launchBrokenLineKernelsOnCPU (via oracle_fit.hpp)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/pixeltrack.h"
#include "oracle.h"
#include "oracle_fit.hpp"

namespace {

// ---------------------------------------------------------------- constants
// gpu_pixel_doublets.jl:20-28
const int kLayerPairs[2 * 19] = {0, 1, 0, 4, 0, 7, 1, 2, 1, 4, 1, 7, 4, 5, 7, 8, 2, 3, 2, 4,
                                 2, 7, 5, 6, 8, 9, 0, 2, 1, 3, 0, 5, 0, 8, 4, 6, 7, 9};
// gpu_pixel_doublets.jl:29-52
const int16_t kPhiCuts[19] = {522, 730, 730, 522, 626, 626, 522, 522, 626, 626,
                              626, 522, 522, 522, 522, 522, 522, 522, 522};
// gpu_pixel_doublets.jl:54-56
const double kMinZ[19] = {-20., 0., -30., -22., 10., -30., -70., -70., -22., 15.,
                          -30, -70., -70., -20., -22., 0, -30., -70., -70.};
const double kMaxZ[19] = {20., 30., 0., 22., 30., -10., 70., 70., 22., 30.,
                          -15., 70., 70., 20., 22., 30., 0., 70., 70.};
const double kMaxR[19] = {20., 9., 9., 20., 7., 7., 5., 5., 20., 6.,
                          6., 5., 5., 20., 20., 9., 9., 9., 9.};

// hist_to_container.jl:81-86 with T=Int16, 128 bins, S=16: ((t >> 9) & 127) + 1
inline int hist_bin(int16_t t) { return (int)(((uint16_t)t) >> 9) + 1; }

// ---------------------------------------------------------------- PreAllocMatrix (simple_matrix.jl:1-58)
template <class T> struct PreAllocMatrix {
    uint32_t nrowmax = 0, ncolmax = 0, ncol = 0;
    std::vector<T> data;
    std::vector<uint32_t> len;
    // (re)allocation keeps old storage: like Matrix{T}(undef, ...) only the
    // column lengths are initialised (simple_matrix.jl:1-11)
    void init(uint32_t nr, uint32_t nc) {
        nrowmax = nr;
        ncolmax = nc;
        ncol = 0;
        if (data.size() < (size_t)nr * nc) data.resize((size_t)nr * nc);
        len.assign(nc, 0);
    }
    // extend!: returns the last column (not failure) when exhausted (simple_matrix.jl:13-21)
    uint32_t extend() {
        if (ncol == ncolmax) return ncolmax;
        ncol += 1;
        return ncol;
    }
    // push!: silently drops when full (simple_matrix.jl:39-49); col is 1-based
    uint32_t push(uint32_t col, T v) {
        uint32_t row = len[col - 1];
        if (row == nrowmax) return row;
        len[col - 1] += 1;
        data[(size_t)(col - 1) * nrowmax + row] = v;
        return row + 1;
    }
    uint32_t length(uint32_t col) const { return len[col - 1]; }
    T at(uint32_t col, uint32_t row1) const { return data[(size_t)(col - 1) * nrowmax + (row1 - 1)]; }
    void reset() {
        ncol = 0;
        std::fill(len.begin(), len.end(), 0u);
    }
};

// GPUCACell (gpu_ca_cell.jl:28-38)
struct Cell {
    uint32_t outer_neighbors;
    uint32_t tracks;
    int32_t doublet_id;
    int16_t layer_pair_id;  // 1-based
    uint16_t used;
    float inner_z, inner_r;
    uint16_t inner_hit_id, outer_hit_id;  // 1-based
};

struct Hits {
    const pt_hits_view *v;
    // 1-based accessors
    float xg(uint32_t i) const { return v->xg[i - 1]; }
    float yg(uint32_t i) const { return v->yg[i - 1]; }
    float zg(uint32_t i) const { return v->zg[i - 1]; }
    float rg(uint32_t i) const { return v->rg[i - 1]; }
    int16_t iphi(uint32_t i) const { return v->iphi[i - 1]; }
    int16_t ysize(uint32_t i) const { return v->ysize[i - 1]; }
    int16_t det(uint32_t i) const { return v->det_ind[i - 1]; }
};

struct State {
    std::vector<Cell> cells;
    uint32_t n_cells = 0;
    uint32_t n_found = 0;
    PreAllocMatrix<uint32_t> ohc;  // OuterHitOfCellVector(128, n_hits)
    PreAllocMatrix<uint32_t> nbr;  // CellNeighborsVector(64, 65536)
    PreAllocMatrix<uint16_t> trk;  // CellTracksVector(64, 65536)
    uint32_t apc[2] = {0, 0};      // device_hit_tuple_counter
    std::vector<uint32_t> tup_off;  // HitContainer off [32769]
    std::vector<uint16_t> tup_bins;  // [163840]
    std::vector<int32_t> quality;    // [32768]
    std::vector<uint8_t> quality_set;
    uint32_t tm_off[9];              // TupleMultiplicity off (8 bins + 1)
    std::vector<uint16_t> tm_bins;   // [MAX_TUPLES = 49152]
    uint32_t overflow = 0, ref_assert = 0;
    uint32_t n_connect_pairs = 0, n_connections_raw = 0;
    uint32_t n_nbr_pushes_dropped = 0;
};

// ------------------------------------------------------------ doublet cuts
// gpu_pixel_doublets.jl:61-65 (Float64 dphi and r2t4, Float32 products)
inline bool pt_cut(const Hits &hh, uint32_t j, int16_t idphi, double r2t4, float ri) {
    float ro = hh.rg(j);
    double dphi = (double)idphi * ((2.0 * oracle::kPi) / 65536.0);
    float rr = ri * ro;
    float d = ro - ri;
    float d2 = d * d;
    return dphi * dphi * (r2t4 - (double)rr) > (double)d2;
}
// gpu_pixel_doublets.jl:69-74
inline bool z0_cut_off(const Hits &hh, uint32_t j, float ri, float zi, int pair, float z0_cut) {
    float zo = hh.zg(j);
    float ro = hh.rg(j);
    float dr = ro - ri;
    return (double)dr > kMaxR[pair] || dr < 0.0f || std::fabs(zi * ro - ri * zo) > z0_cut * dr;
}
// gpu_pixel_doublets.jl:79-88 ; me_s is Int (-1) or Int16
inline bool z_size_cut(const Hits &hh, uint32_t j, int outer, int inner, int me_s, float zi, float ri,
                       float dz_dr_fact, uint32_t &ref_assert) {
    const bool only_barrel = outer < 4;
    int so = hh.ysize(j);
    int dy = (inner == 0) ? 28 : 20;
    float zo = hh.zg(j);
    float ro = hh.rg(j);
    if (only_barrel) return me_s > 0 && so > 0 && std::abs(so - me_s) > dy;
    if (!(inner < 4 && me_s > 0)) return false;
    float q = std::fabs((zi - zo) / (ri - ro)) * dz_dr_fact;
    double t = std::trunc((double)q + 0.5);
    if (!std::isfinite(t) || std::fabs(t) > 9.0e15) {
        ref_assert |= PT_REF_SIZE_CUT_INF;  // Int(Inf/NaN) throws InexactError
        return true;
    }
    long long iv = (long long)t;
    return std::llabs((long long)me_s - iv) > 20;
}

// gpu_pixel_doublets.jl:121-360
void doublets_from_histo(const pt_params &P, const Hits &hh, State &st, int n_pairs) {
    const pt_hits_view &v = *hh.v;
    const uint32_t *offsets = v.layer_start;
    auto layer_size = [&](int L) -> uint32_t { return offsets[L + 1] - offsets[L]; };
    const int min_y_size_B1 = 36, min_y_size_B2 = 28;
    const float dz_dr_fact = (float)(8 * 0.0285 / 0.015);
    uint32_t cum[20];
    cum[0] = layer_size(kLayerPairs[0]);
    for (int i = 1; i < n_pairs; ++i) cum[i] = cum[i - 1] + layer_size(kLayerPairs[2 * i]);
    const uint32_t n_tot = cum[n_pairs - 1];
    const float z0_cut = 12.0f;
    const double hard_pt_cut = 0.5;
    const double min_radius = hard_pt_cut * 87.78;
    const double min_radius_2T4 = 4. * min_radius * min_radius;
    int p = 0;  // 0-based pair index
    for (uint32_t j = 0; j < n_tot; ++j) {
        while (j >= cum[p]) ++p;
        const int inner = kLayerPairs[2 * p], outer = kLayerPairs[2 * p + 1];
        const int h_off = 128 * outer;
        uint32_t i = (p == 0) ? j : j - cum[p - 1];
        i += offsets[inner];
        i += 1;  // 1-based
        const int16_t mi = hh.det(i);
        if (mi > 2000) continue;
        const float me_z = hh.zg(i);
        if ((double)me_z < kMinZ[p] || (double)me_z > kMaxZ[p]) continue;
        int me_s = -1;
        if (P.do_cluster_cut) {
            if (inner == 0 && !(mi < 96)) st.ref_assert |= PT_REF_BAD_MODULE;
            bool is_outer_ladder = P.ideal_conditions ? true : (((mi / 8) % 2) == 0);
            me_s = (inner > 0 || is_outer_ladder) ? (int)hh.ysize(i) : -1;
            if (inner == 0 && outer > 3 && me_s > 0 && me_s < min_y_size_B1) continue;
            if (inner == 1 && outer > 3 && me_s > 0 && me_s < min_y_size_B2) continue;
        }
        const int16_t me_p = hh.iphi(i);
        const float me_r = hh.rg(i);
        const int16_t cut = kPhiCuts[p];
        int kl = hist_bin((int16_t)(me_p - cut));
        int kh = hist_bin((int16_t)(me_p + cut)) + 1;
        if (kh == 129) kh = 1;
        int cb = kl;
        while (cb != kh) {
            const uint32_t pb = v.hist_off[cb + h_off - 1] + 1;
            const uint32_t e = v.hist_off[cb + h_off] + 1;
            for (uint32_t pos = pb; pos < e; ++pos) {
                const uint32_t oi = v.hist_bins[pos - 1];
                if (!(oi > offsets[outer] && oi <= offsets[outer + 1])) st.ref_assert |= PT_REF_HIST_LAYER;
                const int16_t mo = hh.det(oi);
                if (mo > 2000) continue;
                if (P.do_z0_cut && z0_cut_off(hh, oi, me_r, me_z, p, z0_cut)) continue;
                const int16_t mo_p = hh.iphi(oi);
                int16_t i_dphi = (int16_t)(mo_p - me_p);
                i_dphi = (int16_t)(i_dphi < 0 ? -i_dphi : i_dphi);  // Int16 abs (wraps at -32768)
                if (i_dphi > cut) continue;
                if (P.do_cluster_cut && z_size_cut(hh, oi, outer, inner, me_s, me_z, me_r, dz_dr_fact, st.ref_assert))
                    continue;
                if (P.do_pt_cut && pt_cut(hh, oi, i_dphi, min_radius_2T4, me_r)) continue;
                // The reference increments, finds the cap exceeded, decrements and
                // breaks out of this bin; no later candidate is ever stored, so
                // counting on without storing is equivalent.
                st.n_found += 1;
                if (st.n_cells + 1 > P.max_num_of_doublets) {
                    st.overflow |= PT_OVF_DOUBLETS;
                    continue;
                }
                st.n_cells += 1;
                Cell c;
                c.outer_neighbors = 1;
                c.tracks = 1;
                c.doublet_id = (int32_t)st.n_cells;
                c.layer_pair_id = (int16_t)(p + 1);
                c.used = 0;
                c.inner_z = hh.zg(i);
                c.inner_r = hh.rg(i);
                c.inner_hit_id = (uint16_t)i;
                c.outer_hit_id = (uint16_t)oi;
                st.cells[st.n_cells - 1] = c;
                if (st.ohc.length(oi) == st.ohc.nrowmax) st.overflow |= PT_OVF_OUTER_HIT_CELLS;
                st.ohc.push(oi, st.n_cells);
            }
            cb += 1;
            if (cb == 129) cb = 1;
        }
    }
}

// gpu_ca_cell.jl:130-136
inline bool are_aligned(float r1, float z1, float ri, float zi, float ro, float zo, float pt_min, float theta_cut) {
    float radius_diff = std::fabs(r1 - ro);
    float d13 = radius_diff * radius_diff + (z1 - zo) * (z1 - zo);
    float p_min = pt_min * std::sqrt(d13);
    float tan_ = std::fabs(z1 * (ri - ro) + zi * (ro - r1) + zo * (r1 - ri));
    return tan_ * p_min <= theta_cut * d13 * radius_diff;
}
// circle_eq.jl:24-44, 89-93 ; gpu_ca_cell.jl:138-154
inline bool dca_cut(float x1, float y1, float x2, float y2, float x3, float y3, float region, float max_curv) {
    bool no_flip = std::fabs(x3 - x1) < std::fabs(y3 - y1);
    float x1p = no_flip ? x1 - x2 : y1 - y2;
    float y1p = no_flip ? y1 - y2 : x1 - x2;
    float d12 = x1p * x1p + y1p * y1p;
    float x3p = no_flip ? x3 - x2 : y3 - y2;
    float y3p = no_flip ? y3 - y2 : x3 - x2;
    float d32 = x3p * x3p + y3p * y3p;
    float num = x1p * y3p - y1p * x3p;
    float det = d12 * y3p - d32 * y1p;
    float st2 = (d12 * x3p - d32 * x1p);
    float seq = det * det + st2 * st2;
    float al2 = 1.0f / std::sqrt(seq);
    float be2 = -st2 * al2;
    float ct = 2.0f * num * al2;
    al2 *= det;
    float m_xp = x2, m_yp = y2;
    float m_c = no_flip ? ct : -ct;
    float m_alpha = no_flip ? al2 : -be2;
    float m_beta = no_flip ? be2 : -al2;
    if (m_c > max_curv) return false;
    float xx = m_c * m_xp + m_alpha;
    float yy = m_c * m_yp + m_beta;
    float dca0 = std::sqrt(xx * xx + yy * yy) - 1.0f;
    return std::fabs(dca0) < region * std::fabs(m_c);
}

// gpu_ca_cell.jl:164-177
void add_outer_neighbor(State &st, uint32_t cell_index, uint32_t t) {
    Cell &c = st.cells[cell_index - 1];
    if (c.outer_neighbors == 1) {
        uint32_t i = st.nbr.extend();  // exhausted columns hand out the last column again
        if (i > 1) {
            c.outer_neighbors = i;
        } else {
            return;
        }
    }
    if (st.nbr.length(c.outer_neighbors) == st.nbr.nrowmax) st.overflow |= PT_OVF_NEIGHBORS;
    st.nbr.push(c.outer_neighbors, t);
}

// ca_hit_ntuplet_generator_kernels_impl.jl:45-114
void kernel_connect(const pt_params &P, const Hits &hh, State &st) {
    const uint32_t last_bpix1_det_index = 96, last_barrel_det_index = 1184;
    for (uint32_t idx = 1; idx <= st.n_cells; ++idx) {
        Cell this_cell = st.cells[idx - 1];
        const uint32_t ih = this_cell.inner_hit_id;
        const uint32_t n_poss = st.ohc.length(ih);
        const float ri = this_cell.inner_r, zi = this_cell.inner_z;
        const float ro = hh.rg(this_cell.outer_hit_id), zo = hh.zg(this_cell.outer_hit_id);
        const bool is_barrel = (int)hh.det(ih) < (int)last_barrel_det_index;
        for (uint32_t j = 1; j <= n_poss; ++j) {
            const uint32_t oidx = st.ohc.at(ih, j);
            Cell other = st.cells[oidx - 1];
            const float r1 = other.inner_r, z1 = other.inner_z;
            st.n_connect_pairs += 1;
            bool aligned = are_aligned(r1, z1, ri, zi, ro, zo, P.pt_min,
                                       is_barrel ? P.ca_theta_cut_barrel : P.ca_theta_cut_forward);
            const float region = ((int)hh.det(other.inner_hit_id) < (int)last_bpix1_det_index)
                                     ? P.dca_cut_inner_triplet
                                     : P.dca_cut_outer_triplet;
            bool cut = dca_cut(hh.xg(other.inner_hit_id), hh.yg(other.inner_hit_id), hh.xg(this_cell.inner_hit_id),
                               hh.yg(this_cell.inner_hit_id), hh.xg(this_cell.outer_hit_id),
                               hh.yg(this_cell.outer_hit_id), region, P.hard_curv_cut);
            if (aligned && cut) {
                st.n_connections_raw += 1;
                add_outer_neighbor(st, oidx, idx);
                if (this_cell.used != 1) {
                    this_cell = st.cells[idx - 1];
                    st.cells[idx - 1].used |= 1;
                    this_cell.used |= 1;
                }
                if (other.used != 1) st.cells[oidx - 1].used |= 1;
            }
        }
    }
}

// gpu_fishbone.jl:4-62
// Planted mistakes in fish_bone, set only by the tests that show the
// fishbone pins catch them (default: none).
struct FishPlant {
    bool no_break = false, no_module_check = false;
};
FishPlant g_fish_plant;

void fish_bone(const Hits &hh, State &st, uint32_t n_hits, bool check_track) {
    float dx[128], dy[128], dz[128], norms[128];
    uint16_t inner_det[128];
    uint32_t cidx[128];
    for (uint32_t idy = 1; idy <= n_hits; ++idy) {
        const uint32_t num = st.ohc.length(idy);
        if (num < 2) continue;
        const Cell &first = st.cells[st.ohc.at(idy, 1) - 1];
        const float xo = hh.xg(first.outer_hit_id), yo = hh.yg(first.outer_hit_id), zo = hh.zg(first.outer_hit_id);
        int curr = 0;
        for (uint32_t ic = 1; ic <= num; ++ic) {
            const uint32_t ci = st.ohc.at(idy, ic);
            const Cell &c = st.cells[ci - 1];
            if (c.used == 0) continue;
            if (check_track && st.trk.length(c.tracks) == 0) continue;
            cidx[curr] = ci;
            inner_det[curr] = (uint16_t)hh.det(c.inner_hit_id);
            dx[curr] = hh.xg(c.inner_hit_id) - xo;
            dy[curr] = hh.yg(c.inner_hit_id) - yo;
            dz[curr] = c.inner_z - zo;
            norms[curr] = dx[curr] * dx[curr] + dy[curr] * dy[curr] + dz[curr] * dz[curr];
            ++curr;
        }
        if (curr < 2) continue;
        for (int ic = 0; ic < curr - 1; ++ic) {
            for (int jc = ic + 1; jc < curr; ++jc) {
                float cos12 = dx[ic] * dx[jc] + dy[ic] * dy[jc] + dz[ic] * dz[jc];
                const bool other_module = g_fish_plant.no_module_check || inner_det[ic] != inner_det[jc];
                if (other_module && cos12 * cos12 >= 0.99999f * norms[ic] * norms[jc]) {
                    if (norms[ic] > norms[jc]) {
                        st.cells[cidx[ic] - 1].doublet_id = -1;
                        if (g_fish_plant.no_break) continue;
                        break;
                    } else {
                        st.cells[cidx[jc] - 1].doublet_id = -1;
                    }
                }
            }
        }
    }
}

// hist_to_container.jl:131-146 ; returns tuple id (1-based) or <= 0 on overflow
int64_t bulk_fill(State &st, const uint16_t *v, uint32_t n) {
    uint32_t c0 = st.apc[0], c1 = st.apc[1];
    st.apc[0] += 1;
    st.apc[1] += n;
    if (st.apc[0] > PT_MAX_TUPLES) {
        st.overflow |= PT_OVF_TUPLES;
        return -(int64_t)c0;
    }
    st.tup_off[st.apc[0] - 1] = c1 + 1;
    for (uint32_t i = 0; i < n; ++i) st.tup_bins[c1 + i] = v[i];
    return st.apc[0];
}

// gpu_ca_cell.jl:179-191
void add_track(State &st, uint32_t cell_index, uint32_t t) {
    Cell &c = st.cells[cell_index - 1];
    if (c.tracks == 1) {
        uint32_t i = st.trk.extend();
        if (i > 1) {
            c.tracks = i;
        } else {
            return;
        }
    }
    if (st.trk.length(c.tracks) == st.trk.nrowmax) st.overflow |= PT_OVF_CELL_TRACKS;
    st.trk.push(c.tracks, (uint16_t)t);
}

// gpu_ca_cell.jl:232-270 ; the Stack holds doublet ids, iterated top -> bottom
void find_ntuplets(const pt_params &P, State &st, const Cell self, std::vector<uint32_t> &stack) {
    stack.push_back((uint32_t)self.doublet_id);
    if (stack.size() > 4) {
        st.ref_assert |= PT_REF_DEPTH;  // @assert length(temp_ntuplet) <= 4
        stack.pop_back();
        return;
    }
    bool last = true;
    const uint32_t col = self.outer_neighbors;
    const uint32_t nn = st.nbr.length(col);
    for (uint32_t k = 1; k <= nn; ++k) {
        const uint32_t other = st.nbr.at(col, k);
        if (st.cells[other - 1].doublet_id < 0) continue;
        last = false;
        if (st.cells[other - 1].inner_hit_id == self.inner_hit_id) st.ref_assert |= PT_REF_DEPTH;
        find_ntuplets(P, st, st.cells[other - 1], stack);
    }
    if (last) {
        if (stack.size() >= P.min_hits_per_ntuplet - 1) {
            uint16_t hits[6] = {0, 0, 0, 0, 0, 0};
            const size_t len = stack.size();
            // iterate the stack from the top (LIFO) writing hits[n_h], n_h decreasing
            size_t n_h = len;
            for (size_t s = len; s-- > 0;) {
                hits[n_h - 1] = st.cells[stack[s] - 1].inner_hit_id;
                n_h -= 1;
            }
            hits[len] = self.outer_hit_id;
            int64_t it = bulk_fill(st, hits, (uint32_t)len + 1);
            if (it > 0) {
                for (size_t s = len; s-- > 0;) add_track(st, stack[s], (uint32_t)it);
                st.quality[it - 1] = PT_QUALITY_BAD;
                st.quality_set[it - 1] = 1;
            }
        }
    }
    stack.pop_back();
}

// ca_hit_ntuplet_generator_kernels_impl.jl:116-132
void kernel_find_ntuplets(const pt_params &P, State &st) {
    std::vector<uint32_t> stack;
    stack.reserve(8);
    for (uint32_t idx = 1; idx <= st.n_cells; ++idx) {
        const Cell this_cell = st.cells[idx - 1];
        if (this_cell.doublet_id < 0) continue;
        int p_id = this_cell.layer_pair_id - 1;
        bool do_it = (P.min_hits_per_ntuplet > 3) ? (p_id < 3) : (p_id < 8 || p_id > 12);
        if (do_it) find_ntuplets(P, st, this_cell, stack);
    }
}

// hist_to_container.jl:150-161
void bulk_finalize_fill(State &st) {
    const uint32_t m = st.apc[0], n = st.apc[1];
    if (m > PT_MAX_TUPLES) {
        st.tup_off[PT_MAX_TUPLES] = st.tup_off[PT_MAX_TUPLES - 1];
        return;
    }
    for (uint32_t i = m + 1; i <= PT_MAX_TUPLES + 1; ++i) st.tup_off[i - 1] = n + 1;
}

inline uint32_t tuple_size(const State &st, uint32_t it) { return st.tup_off[it] - st.tup_off[it - 1]; }

// ca_hit_ntuplet_generator_kernels_impl.jl:143-166
void kernel_early_duplicate_remover(State &st) {
    if (st.n_cells == 0) st.ref_assert |= PT_REF_NO_CELLS;
    for (uint32_t idx = 1; idx <= st.n_cells; ++idx) {
        const uint32_t col = st.cells[idx - 1].tracks;
        const uint32_t nt = st.trk.length(col);
        if (nt < 2) continue;
        uint32_t maxh = 0;
        for (uint32_t k = 1; k <= nt; ++k) maxh = std::max(maxh, tuple_size(st, st.trk.at(col, k)));
        for (uint32_t k = 1; k <= nt; ++k) {
            uint32_t it = st.trk.at(col, k);
            if (tuple_size(st, it) != maxh) st.quality[it - 1] = PT_QUALITY_DUP;
        }
    }
}

// ca_hit_ntuplet_generator_kernels_impl.jl:174-211 + finalize! (hist_to_container.jl:201-205)
void multiplicity(State &st) {
    for (int b = 0; b < 9; ++b) st.tm_off[b] = 0;
    for (uint32_t it = 1; it <= PT_MAX_TUPLES; ++it) {
        uint32_t nh = tuple_size(st, it);
        if (nh < 3) continue;
        if (st.quality[it - 1] == PT_QUALITY_DUP) continue;
        if (nh >= 8) {
            st.ref_assert |= PT_REF_MULT;
            continue;
        }
        st.tm_off[nh - 1] += 1;
    }
    for (int b = 1; b < 9; ++b) st.tm_off[b] += st.tm_off[b - 1];
    for (uint32_t it = 1; it <= PT_MAX_TUPLES; ++it) {
        uint32_t nh = tuple_size(st, it);
        if (nh < 3) continue;
        if (st.quality[it - 1] == PT_QUALITY_DUP) continue;
        if (nh >= 8) continue;
        uint32_t w = st.tm_off[nh - 1];
        st.tm_off[nh - 1] -= 1;
        st.tm_bins[w - 1] = (uint16_t)it;
    }
}

// fit of one bucket (broken_line_fit_on_gpu.jl:24-163), N hits used
template <int N>
void fit_bucket(const Hits &hh, const float *frames, uint32_t nHits, double B, const State &st, pt_track_soa *out,
                uint32_t &n_fitted, uint32_t &ref_assert) {
    const uint32_t cnt = st.tm_off[nHits] - st.tm_off[nHits - 1];
    for (uint32_t local = 1; local <= PT_MAX_FITS; ++local) {
        if (local >= cnt + 1) break;
        const uint32_t tkid = st.tm_bins[st.tm_off[nHits - 1] + local - 1];
        double hits[3][N];
        double ge[6][N];
        const uint32_t start = st.tup_off[tkid - 1];
        for (int i = 0; i < N; ++i) {
            const uint32_t hit = st.tup_bins[start - 1 + i];
            const int16_t det = hh.det(hit);
            float g[6];
            oracle::to_global_err(frames + 9 * (size_t)det, hh.v->xerr[hit - 1], 0.0f, hh.v->yerr[hit - 1], g);
            hits[0][i] = hh.xg(hit);
            hits[1][i] = hh.yg(hit);
            hits[2][i] = hh.zg(hit);
            for (int c = 0; c < 6; ++c) ge[c][i] = g[c];
        }
        oracle::TrackOut<double> o;
        oracle::fit_track<double, N>(hits, ge, B, o);
        if (o.flags & oracle::kFlagFastNaN) ref_assert |= PT_REF_FIT_NAN;
        if (o.flags & oracle::kFlagCharge) ref_assert |= PT_REF_FIT_CHARGE;
        if (o.flags & oracle::kFlagSqrtDomain) ref_assert |= PT_REF_SQRT_DOMAIN;
        const size_t row = tkid - 1;
        for (int k = 0; k < 5; ++k) out->state[k * (size_t)PT_MAX_TUPLES + row] = (float)o.state[k];
        for (int k = 0; k < 15; ++k) out->covariance[k * (size_t)PT_MAX_TUPLES + row] = (float)o.cov15[k];
        out->pt[row] = (float)o.pt;
        out->eta[row] = (float)o.eta;
        out->chi2[row] = (float)o.chi2;
        n_fitted += 1;
    }
}

// ca_hit_ntuplet_generator_kernels_impl.jl:213-247
void kernel_classify_tracks(const pt_params &P, const State &st, pt_track_soa *out) {
    for (uint32_t it = 1; it <= PT_MAX_TUPLES; ++it) {
        const uint32_t nh = tuple_size(st, it);
        if (nh == 0) break;
        if (out->quality[it - 1] == PT_QUALITY_DUP) continue;
        if (nh < 3) continue;
        bool is_nan = false;
        for (int i = 0; i < 5; ++i) is_nan |= std::isnan(out->state[i * (size_t)PT_MAX_TUPLES + it - 1]);
        if (is_nan) continue;
        const float ptv = out->pt[it - 1];
        // Julia min(Float32, Float32): NaN propagates
        const float pt = std::isnan(ptv) ? ptv : (P.chi2_max_pt < ptv ? P.chi2_max_pt : ptv);
        const float chi2_cut =
            P.chi2_scale *
            (P.chi2_coeff[0] + pt * (P.chi2_coeff[1] + pt * (P.chi2_coeff[2] + pt * P.chi2_coeff[3])));
        if (3.0f * out->chi2[it - 1] >= chi2_cut) continue;
        const bool quad = nh > 3;
        const float max_tip = quad ? P.quadruplet_max_tip : P.triplet_max_tip;
        const float min_pt = quad ? P.quadruplet_min_pt : P.triplet_min_pt;
        const float max_zip = quad ? P.quadruplet_max_zip : P.triplet_max_zip;
        const float tip = out->state[1 * (size_t)PT_MAX_TUPLES + it - 1];
        const float zip = out->state[4 * (size_t)PT_MAX_TUPLES + it - 1];
        if (std::fabs(tip) < max_tip && out->pt[it - 1] > min_pt && std::fabs(zip) < max_zip)
            out->quality[it - 1] = PT_QUALITY_LOOSE;
    }
}

// ca_hit_ntuplet_generator_kernels_impl.jl:250-279
void kernel_fast_duplicate_remover(const State &st, pt_track_soa *out) {
    for (uint32_t idx = 1; idx <= st.n_cells; ++idx) {
        const uint32_t col = st.cells[idx - 1].tracks;
        const uint32_t nt = st.trk.length(col);
        if (nt < 2) continue;
        float mc = 10000.0f;
        uint32_t im = 60000;
        for (uint32_t k = 1; k <= nt; ++k) {
            const uint32_t it = st.trk.at(col, k);
            const float sc = std::fabs(out->state[1 * (size_t)PT_MAX_TUPLES + it - 1]);
            if (out->quality[it - 1] == PT_QUALITY_LOOSE && sc < mc) {
                mc = sc;
                im = it;
            }
        }
        for (uint32_t k = 1; k <= nt; ++k) {
            const uint32_t it = st.trk.at(col, k);
            if (out->quality[it - 1] != PT_QUALITY_BAD && it != im) out->quality[it - 1] = PT_QUALITY_DUP;
        }
    }
}

// ca_hit_ntuplet_generator_kernels_impl.jl:281-362 with HitToTuple = OneToManyAssoc{UInt16,49152,196608}
void hit_to_tuple_and_triplet_cleaner(const State &st, pt_track_soa *out, uint32_t &overflow) {
    const uint32_t NB = PT_MAX_HITS;  // bins
    std::vector<uint32_t> off(NB + 1, 0);
    std::vector<uint16_t> bins(4 * 49152);
    auto loop = [&](bool fill) {
        for (uint32_t idx = 1; idx <= PT_MAX_TUPLES; ++idx) {
            if (tuple_size(st, idx) == 0) break;
            if (out->quality[idx - 1] != PT_QUALITY_LOOSE) continue;
            // for h in begin_h(tuples,idx)-1 : end_h(tuples,idx)-2  (begin_h = off[idx]+1)
            const uint32_t hb = st.tup_off[idx - 1] + 1 - 1;
            const uint32_t he = st.tup_off[idx] + 1 - 2;
            for (uint32_t h = hb; h <= he; ++h) {
                const uint32_t b = st.tup_bins[h - 1];  // 1-based hit id used as bin
                if (!fill) {
                    off[b - 1] += 1;
                } else {
                    uint32_t w = off[b - 1];
                    off[b - 1] -= 1;
                    if (w == 0 || w > bins.size()) {
                        overflow |= PT_OVF_HIT_TO_TUPLE;
                        continue;
                    }
                    bins[w - 1] = (uint16_t)idx;
                }
            }
        }
    };
    loop(false);
    for (uint32_t b = 1; b <= NB; ++b) off[b] += off[b - 1];
    loop(true);
    for (uint32_t idx = 1; idx <= NB; ++idx) {
        const uint32_t b0 = off[idx - 1] + 1, b1 = off[idx] + 1;  // begin_h, end_h
        if (off[idx] - off[idx - 1] < 2) continue;
        float mc = 10000.0f;
        uint32_t im = 60000;
        uint32_t maxNh = 0;
        for (uint32_t p = b0; p < b1; ++p) maxNh = std::max(maxNh, tuple_size(st, bins[p - 1]));
        for (uint32_t p = b0; p < b1; ++p) {
            const uint32_t it = bins[p - 1];
            if (maxNh != tuple_size(st, it)) out->quality[it - 1] = PT_QUALITY_DUP;
        }
        if (maxNh > 3) continue;
        for (uint32_t p = b0; p < b1; ++p) {
            const uint32_t it = bins[p - 1];
            const float tip = std::fabs(out->state[1 * (size_t)PT_MAX_TUPLES + it - 1]);
            if (out->quality[it - 1] != PT_QUALITY_BAD && tip < mc) {
                mc = tip;
                im = it;
            }
        }
        for (uint32_t p = b0; p < b1; ++p) {
            const uint32_t it = bins[p - 1];
            if (out->quality[it - 1] != PT_QUALITY_BAD && it != im) out->quality[it - 1] = PT_QUALITY_DUP;
        }
    }
}

}  // namespace

extern "C" {

void oracle_params_default(pt_params *p) {
    std::memset(p, 0, sizeof(*p));
    p->min_hits_per_ntuplet = 3;
    p->max_num_of_doublets = 458752;
    p->use_riemann_fit = 0;
    p->fit_5_as_4 = 1;
    p->include_jumping_forward_doublets = 1;
    p->early_fish_bone = 1;
    p->late_fish_bone = 0;
    p->ideal_conditions = 1;
    p->do_stats = 0;
    p->do_cluster_cut = 1;
    p->do_z0_cut = 1;
    p->do_pt_cut = 1;
    p->pt_min = (float)0.899999976158;
    p->ca_theta_cut_barrel = (float)0.00200000009499;
    p->ca_theta_cut_forward = (float)0.00300000002608;
    p->hard_curv_cut = (float)0.0328407224959;
    p->dca_cut_inner_triplet = (float)0.15000000596;
    p->dca_cut_outer_triplet = (float)0.25;
    p->chi2_coeff[0] = (float)0.68177776;
    p->chi2_coeff[1] = (float)0.74609577;
    p->chi2_coeff[2] = (float)-0.08035491;
    p->chi2_coeff[3] = (float)0.00315399;
    p->chi2_max_pt = 10.0f;
    p->chi2_scale = 30.0f;
    p->triplet_max_tip = 0.3f;
    p->triplet_min_pt = 0.5f;
    p->triplet_max_zip = 12.0f;
    p->quadruplet_max_tip = 0.5f;
    p->quadruplet_min_pt = 0.3f;
    p->quadruplet_max_zip = 12.0f;
}

void oracle_plant_fishbone(int no_break, int no_module_check) {
    g_fish_plant.no_break = no_break != 0;
    g_fish_plant.no_module_check = no_module_check != 0;
}

int oracle_make_tuples(const pt_params *P, const pt_hits_view *hv, const float *frames, double b_field,
                       pt_track_soa *out, pt_counters *cnt, pt_ca_debug *dbg, int stop_after) {
    if (!P || !hv || !out) return PT_ERR_ARG;
    // switches on which the reference throws before producing tracks:
    // late_fish_bone (empty() on a UInt32, gpu_fishbone.jl:29), do_stats
    // (kernel_mark_used is not defined, ca_hit_ntuplet_generator_kernels.jl:182-184),
    // fit_5_as_4 = false (log_file is not defined, broken_line_fit_on_gpu.jl:195-200).
    // use_riemann_fit is never read: make_tuples runs the broken-line fit
    // whatever its value (ca_hit_ntuplet_generator.jl:83-93)
    if (P->late_fish_bone || P->do_stats || !P->fit_5_as_4) return PT_ERR_UNSUPPORTED;
    // per-thread scratch reused across events, like CACHED_KERNELS (a TaskLocalValue,
    // ca_hit_ntuplet_generator.jl:9) reset by resetCAHitNTupletGeneratorKernels
    thread_local State st_tls;
    State &st = st_tls;
    st.n_cells = 0;
    st.n_found = 0;
    st.apc[0] = st.apc[1] = 0;
    st.overflow = st.ref_assert = 0;
    st.n_connect_pairs = st.n_connections_raw = st.n_nbr_pushes_dropped = 0;
    Hits hh{hv};
    const uint32_t n_hits = hv->n_hits;
    std::memset(cnt, 0, sizeof(*cnt));
    cnt->n_hits = n_hits;
    if (n_hits > PT_MAX_HITS) st.ref_assert |= PT_REF_TOO_MANY_HITS;
    // TrackSOA(): off zeroed by zero(tuples_d)
    st.tup_off.assign(PT_MAX_TUPLES + 1, 0);
    if (st.tup_bins.size() < PT_MAX_TUPLE_HITS) st.tup_bins.resize(PT_MAX_TUPLE_HITS);
    if (st.quality.size() < PT_MAX_TUPLES) st.quality.resize(PT_MAX_TUPLES);
    st.quality_set.assign(PT_MAX_TUPLES, 0);
    if (st.tm_bins.size() < 48 * 1024) st.tm_bins.resize(48 * 1024);
    if (st.cells.size() < P->max_num_of_doublets) st.cells.resize(P->max_num_of_doublets);
    st.ohc.init(PT_MAX_CELLS_PER_HIT, std::max<uint32_t>(n_hits, 1));
    if (st.nbr.ncolmax == 0) st.nbr.init(PT_MAX_NEIGHBORS, PT_MAX_ACTIVE_LISTS);
    if (st.trk.ncolmax == 0) st.trk.init(PT_MAX_CELL_TRACKS, PT_MAX_ACTIVE_LISTS);
    st.nbr.reset();
    st.trk.reset();
    // init_doublets (gpu_pixel_doublets.jl:91-106)
    st.nbr.extend();
    st.trk.extend();

    int n_actual_pairs = 19;
    if (!P->include_jumping_forward_doublets) n_actual_pairs = 15;
    if (P->min_hits_per_ntuplet > 3) n_actual_pairs = 13;
    uint32_t n_inner = 0;
    for (int p = 0; p < n_actual_pairs; ++p)
        n_inner += hv->layer_start[kLayerPairs[2 * p] + 1] - hv->layer_start[kLayerPairs[2 * p]];
    cnt->n_inner_slots = n_inner;
    if (n_hits > 0) doublets_from_histo(*P, hh, st, n_actual_pairs);
    if (stop_after != 1) {
        kernel_connect(*P, hh, st);
        if (stop_after != 2) {
            if (n_hits > 1 && P->early_fish_bone) fish_bone(hh, st, n_hits, false);
            if (stop_after != 3) {
                kernel_find_ntuplets(*P, st);
                bulk_finalize_fill(st);
                kernel_early_duplicate_remover(st);
                multiplicity(st);
            }
        }
    }
    // neighbour / track list exhaustion (extend! returned the last column again)
    if (st.nbr.ncol == st.nbr.ncolmax) {
        uint32_t users = 0;
        for (uint32_t i = 0; i < st.n_cells; ++i) users += (st.cells[i].outer_neighbors == st.nbr.ncolmax);
        if (users > 1) st.overflow |= PT_OVF_NEIGHBOR_LISTS;
    }
    if (st.trk.ncol == st.trk.ncolmax) {
        uint32_t users = 0;
        for (uint32_t i = 0; i < st.n_cells; ++i) users += (st.cells[i].tracks == st.trk.ncolmax);
        if (users > 1) st.overflow |= PT_OVF_TRACK_LISTS;
    }
    const bool write_out = stop_after <= 0 || stop_after >= 4;
    const bool do_fit = stop_after <= 0 || stop_after == 5;
    const bool do_classify = stop_after <= 0;
    if (write_out) {
        // output hit_indices (the TrackSOA container) and quality
        for (uint32_t i = 0; i <= PT_MAX_TUPLES; ++i) out->hit_off[i] = st.tup_off[i];
        const uint32_t nbins_written = std::min<uint32_t>(st.apc[1], PT_MAX_TUPLE_HITS);
        for (uint32_t i = 0; i < nbins_written; ++i) out->hit_bins[i] = st.tup_bins[i];
        for (uint32_t i = 0; i < PT_MAX_TUPLES; ++i)
            if (st.quality_set[i]) out->quality[i] = st.quality[i];
        // fill_hit_det_indices (ca_hit_ntuplet_generator_kernels_impl.jl:25-38)
        for (uint32_t i = 0; i <= PT_MAX_TUPLES; ++i) out->det_off[i] = st.tup_off[i];
        const uint32_t sz = st.tup_off[PT_MAX_TUPLES] - 1;
        for (uint32_t i = 1; i <= sz && i <= PT_MAX_TUPLE_HITS; ++i)
            out->det_bins[i - 1] = (uint16_t)hh.det(st.tup_bins[i - 1]);
        // fit (launchBrokenLineKernelsOnCPU, broken_line_fit_on_gpu.jl:166-203)
        const double B = (double)(float)b_field;
        uint32_t n_fitted = 0;
        if (frames && do_fit) {
            fit_bucket<3>(hh, frames, 3, B, st, out, n_fitted, st.ref_assert);
            fit_bucket<4>(hh, frames, 4, B, st, out, n_fitted, st.ref_assert);
            fit_bucket<4>(hh, frames, 5, B, st, out, n_fitted, st.ref_assert);
        }
        cnt->n_fitted = n_fitted;
        // classify_tuples (ca_hit_ntuplet_generator_kernels.jl:222-231)
        if (frames && do_fit && do_classify) {
            kernel_classify_tracks(*P, st, out);
            kernel_fast_duplicate_remover(st, out);
            hit_to_tuple_and_triplet_cleaner(st, out, st.overflow);
        }
        uint32_t n_loose = 0;
        for (uint32_t i = 0; i < std::min<uint32_t>(st.apc[0], PT_MAX_TUPLES); ++i)
            n_loose += (out->quality[i] == PT_QUALITY_LOOSE);
        cnt->n_loose = n_loose;
    }
    for (int b = 3; b < 8; ++b) {
        uint32_t c = st.tm_off[b] - st.tm_off[b - 1];
        cnt->n_by_len[b] = c;
        if (c > PT_MAX_FITS && b <= 5) st.overflow |= PT_OVF_FITS;
    }
    cnt->n_cells = st.n_cells;
    cnt->n_cells_found = st.n_found;
    uint32_t nconn = 0;
    for (uint32_t i = 0; i < st.n_cells; ++i) {
        const uint32_t col = st.cells[i].outer_neighbors;
        if (col != 1) nconn += st.nbr.length(col);
    }
    // count each list once even if shared
    if (st.overflow & PT_OVF_NEIGHBOR_LISTS) {
        nconn = 0;
        for (uint32_t c = 2; c <= st.nbr.ncol; ++c) nconn += st.nbr.length(c);
    }
    cnt->n_connections = nconn;
    cnt->n_connect_pairs = st.n_connect_pairs;
    uint32_t nk = 0;
    for (uint32_t i = 0; i < st.n_cells; ++i) nk += (st.cells[i].doublet_id < 0);
    cnt->n_killed = nk;
    cnt->n_tuples = st.apc[0];
    cnt->n_tuple_hits = st.apc[1];
    cnt->overflow = st.overflow;
    cnt->ref_assert = st.ref_assert;

    if (dbg) {
        dbg->n_cells = st.n_cells;
        const uint32_t nc = std::min(st.n_cells, dbg->cap_cells);
        for (uint32_t i = 0; i < nc; ++i) {
            const Cell &c = st.cells[i];
            if (dbg->cell_inner) dbg->cell_inner[i] = c.inner_hit_id;
            if (dbg->cell_outer) dbg->cell_outer[i] = c.outer_hit_id;
            if (dbg->cell_pair) dbg->cell_pair[i] = (uint8_t)(c.layer_pair_id - 1);
            if (dbg->cell_doublet_id) dbg->cell_doublet_id[i] = c.doublet_id;
            if (dbg->cell_used) dbg->cell_used[i] = (uint8_t)c.used;
        }
        if (dbg->nbr_off && dbg->nbr) {
            uint32_t pos = 0;
            dbg->nbr_off[0] = 0;
            for (uint32_t i = 0; i < nc; ++i) {
                const uint32_t col = st.cells[i].outer_neighbors;
                const uint32_t len = (col == 1) ? 0 : st.nbr.length(col);
                for (uint32_t k = 1; k <= len && pos < dbg->cap_nbr; ++k) dbg->nbr[pos++] = st.nbr.at(col, k);
                dbg->nbr_off[i + 1] = pos;
            }
        }
        if (dbg->trk_off && dbg->trk) {
            uint32_t pos = 0;
            dbg->trk_off[0] = 0;
            for (uint32_t i = 0; i < nc; ++i) {
                const uint32_t col = st.cells[i].tracks;
                const uint32_t len = (col == 1) ? 0 : st.trk.length(col);
                for (uint32_t k = 1; k <= len && pos < dbg->cap_trk; ++k) dbg->trk[pos++] = st.trk.at(col, k);
                dbg->trk_off[i + 1] = pos;
            }
        }
        if (dbg->ohc_off && dbg->ohc) {
            uint32_t pos = 0;
            dbg->ohc_off[0] = 0;
            for (uint32_t h = 1; h <= n_hits; ++h) {
                const uint32_t len = st.ohc.length(h);
                for (uint32_t k = 1; k <= len && pos < dbg->cap_ohc; ++k) dbg->ohc[pos++] = st.ohc.at(h, k);
                dbg->ohc_off[h] = pos;
            }
        }
    }
    return PT_OK;
}

// fill_many_from_vector (hist_to_container.jl:217-270) with the searchsortedfirst
// layer lookup, including its behaviour on empty layers.
int oracle_build_phi_hist(const int16_t *iphi, uint32_t n_hits, const uint32_t *layer_start, uint32_t *off,
                          uint16_t *bins) {
    const uint32_t nh = 10;
    for (uint32_t b = 0; b < PT_HIST_NOFF; ++b) off[b] = 0;
    auto hist_of = [&](uint32_t i) -> int {
        // searchsortedfirst(offsets, i): first 1-based k with offsets[k] >= i
        uint32_t k = 1;
        while (k <= nh + 1 && layer_start[k - 1] < i) ++k;
        if (k <= nh + 1 && layer_start[k - 1] == i) k += 1;
        return (int)k - 2;
    };
    for (uint32_t i = 0; i < layer_start[nh]; ++i) {
        int ih = hist_of(i);
        if (ih < 0 || ih >= (int)nh) return PT_ERR_ARG;
        off[hist_bin(iphi[i]) + 128 * ih - 1] += 1;
    }
    for (uint32_t b = 1; b < PT_HIST_NOFF; ++b) off[b] += off[b - 1];
    for (uint32_t i = 0; i < layer_start[nh]; ++i) {
        int ih = hist_of(i);
        uint32_t b = hist_bin(iphi[i]) + 128 * ih;
        uint32_t w = off[b - 1];
        off[b - 1] -= 1;
        bins[w - 1] = (uint16_t)(i + 1);
    }
    (void)n_hits;
    return PT_OK;
}

// load_tracks (gpu_vertex_finder.jl:51-82): n_hits_track is the size of the
// det_indices bin (PixelTrackHeterogeneous.jl:62); the loop ends at the first
// row without hits.
int oracle_load_tracks(const pt_track_soa *t, float pt_min, uint32_t *n, uint16_t *index_track, float *zt,
                       float *ezt2, float *ptt2) {
    const size_t S = PT_MAX_TUPLES;
    uint32_t nt = 0;
    for (uint32_t idx = 1; idx <= S; ++idx) {
        const uint32_t n_hits = t->det_off[idx] - t->det_off[idx - 1];
        if (n_hits == 0) break;
        if (n_hits < 4) continue;
        if (t->quality[idx - 1] != PT_QUALITY_LOOSE) continue;
        const float pt = t->pt[idx - 1];
        if (pt < pt_min) continue;
        nt += 1;
        index_track[nt - 1] = (uint16_t)idx;
        zt[nt - 1] = t->state[4 * S + idx - 1];
        ezt2[nt - 1] = t->covariance[14 * S + idx - 1];
        ptt2[nt - 1] = pt * pt;
    }
    *n = nt;
    return PT_OK;
}

}  // extern "C"
