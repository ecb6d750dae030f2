// Scene construction and plan generation for the MI355X solver (see scene_model.h).
#include "scene_model.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>

using OpenPSTD::Kernel::PSTDSolverError;

namespace pstdamd {

namespace {

const float kEpsilon = 1e-5f;  // kernel/core/kernel_functions.h:65

int next_pow2(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

struct Interval {
    int a = 0, b = 0;  // [a, b)
    bool empty() const { return b <= a; }
};
Interval intersect(Interval x, Interval y) { return {std::max(x.a, y.a), std::min(x.b, y.b)}; }

Side opposite(Side s) {
    switch (s) {
        case kLeft: return kRight;
        case kRight: return kLeft;
        case kTop: return kBottom;
        default: return kTop;
    }
}

// Reflection/transmission stencil of get_rho_array (kernel/core/kernel_functions.cpp:37-64),
// assembled in double.  Returns the four coefficients a segment uses.
void stencil(double rho1, double rho_self, double rho2, int field, float out[4]) {
    double zn1 = rho1 / rho_self, inv1 = rho_self / rho1;
    double zn2 = rho2 / rho_self, inv2 = rho_self / rho2;
    double rl1 = (zn1 - 1) / (zn1 + 1), tl1 = 2 * zn1 / (zn1 + 1), tl2 = 2 * inv1 / (inv1 + 1);
    double rr1 = (zn2 - 1) / (zn2 + 1), tr1 = 2 * zn2 / (zn2 + 1), tr2 = 2 * inv2 / (inv2 + 1);
    if (field == 0) {  // pressure rows: R(2,1)=tlw2 R(0,0)=rlw1 R(3,1)=trw2 R(1,0)=rrw1
        out[0] = (float)tl2;
        out[1] = (float)rl1;
        out[2] = (float)tr2;
        out[3] = (float)rr1;
    } else {  // velocity rows: V(2,1)=tlw1 V(0,0)=-rlw1 V(3,1)=trw1 V(1,0)=-rrw1
        out[0] = (float)tl1;
        out[1] = (float)-rl1;
        out[2] = (float)tr1;
        out[3] = (float)-rr1;
    }
}

}  // namespace

int SceneModel::list_index_of_pool(int pi) const {
    for (int i = 0; i < (int)list_.size(); ++i)
        if (list_[i] == pi) return i;
    return -1;
}

int SceneModel::make_domain(float alpha, int tlx, int tly, int nx, int ny, bool pml, int pml_for,
                            const EdgeParam *edges) {
    DomainModel d;
    d.id = new_id();
    d.tlx = tlx;
    d.tly = tly;
    d.nx = nx;
    d.ny = ny;
    d.alpha = alpha;
    // Domain::initialize_domain (kernel/core/Domain.cpp:65-75), float as in openPSTD
    d.impedance = -((std::sqrt(1 - alpha) + 1) / (std::sqrt(1 - alpha) - 1));
    d.is_pml = pml;
    if (pml) d.pml_for.push_back(pml_for);
    d.rho = d.rigid() ? 1e30f : settings_.rho * d.impedance;
    d.is_secondary = pml && pool_[pml_for].is_pml;
    if (edges) {
        d.has_edges = true;
        for (int s = 0; s < 4; ++s) d.edge[s] = edges[s];
    }
    if (nx <= 0 || ny <= 0) throw PSTDSolverError("domain with an empty extent");
    pool_.push_back(std::move(d));
    return (int)pool_.size() - 1;
}

bool SceneModel::is_neighbour_of(int a, int b) const {
    for (int s = 0; s < 4; ++s)
        for (int n : pool_[a].nb[s])
            if (n == b) return true;
    return false;
}

bool SceneModel::is_pml_for(int a, int b) const {
    for (int p : pool_[a].pml_for)
        if (p == b) return true;
    return false;
}

// Scene::add_domain neighbour rules (kernel/core/Scene.cpp:306-379).
void SceneModel::add_domain(int pi) {
    DomainModel &d = pool_[pi];
    for (int oi : list_) {
        DomainModel &o = pool_[oi];
        if (d.is_secondary && o.is_secondary) continue;
        if ((d.is_secondary && !o.is_pml) || (o.is_secondary && !d.is_pml)) continue;
        if (d.is_pml && o.is_pml) {
            bool o_for_d = is_pml_for(oi, pi), d_for_o = is_pml_for(pi, oi);
            bool parent_link = (d.is_secondary && d_for_o) || (o.is_secondary && o_for_d);
            if (!parent_link) {
                if (o.pml_for.size() != 1 || d.pml_for.size() != 1) continue;
                if (!is_neighbour_of(o.pml_for[0], d.pml_for[0])) continue;
                if (!is_neighbour_of(d.pml_for[0], o.pml_for[0])) continue;
            }
        }
        Side orient;
        if (d.brx() == o.tlx) orient = kRight;
        else if (d.tlx == o.brx()) orient = kLeft;
        else if (d.bry() == o.tly) orient = kBottom;
        else if (d.tly == o.bry()) orient = kTop;
        else continue;
        bool o_foreign = o.is_pml && !d.is_pml && !is_pml_for(oi, pi);
        bool d_foreign = d.is_pml && !o.is_pml && !is_pml_for(pi, oi);
        if (o_foreign || d_foreign) continue;
        Interval own = (orient == kLeft || orient == kRight) ? Interval{d.tly, d.bry()} : Interval{d.tlx, d.brx()};
        Interval oth = (orient == kLeft || orient == kRight) ? Interval{o.tly, o.bry()} : Interval{o.tlx, o.brx()};
        if (intersect(own, oth).empty()) continue;
        d.nb[orient].push_back(oi);
        o.nb[opposite(orient)].push_back(pi);
    }
    list_.push_back(pi);
}

// Scene::add_pml_domains (kernel/core/Scene.cpp:41-225) with
// Domain::get_vacant_range (kernel/core/Domain.cpp:543-584).
void SceneModel::add_pml_domains() {
    const int N = settings_.pml_cells;
    const Side order[4] = {kLeft, kTop, kRight, kBottom};
    std::vector<int> first_order;
    std::vector<std::pair<int, Side>> second_order;  // creation order
    std::vector<int> users = list_;
    for (int ui : users) {
        if (pool_[ui].is_pml) continue;
        for (int i = 0; i < 4; ++i) {
            const Side side = order[i];
            const bool xdir = side == kLeft || side == kRight;
            // vacant part of the edge
            Interval edge = xdir ? Interval{pool_[ui].tly, pool_[ui].bry()} : Interval{pool_[ui].tlx, pool_[ui].brx()};
            std::vector<char> covered(edge.b - edge.a, 0);
            for (int n : pool_[ui].nb[side]) {
                Interval r = xdir ? Interval{pool_[n].tly, pool_[n].bry()} : Interval{pool_[n].tlx, pool_[n].brx()};
                Interval c = intersect(edge, r);
                for (int v = c.a; v < c.b; ++v) covered[v - edge.a] = 1;
            }
            std::vector<Interval> vacant;
            for (int v = edge.a; v < edge.b; ++v) {
                if (covered[v - edge.a]) continue;
                if (vacant.empty() || vacant.back().b != v) vacant.push_back({v, v + 1});
                else vacant.back().b = v + 1;
            }
            if (vacant.size() > 1)
                throw PSTDSolverError("an edge with more than one vacant segment is not supported "
                                      "(openPSTD's get_vacant_range corrupts that case)");
            for (const Interval &vr : vacant) {
                const DomainModel &u = pool_[ui];
                float alpha = u.edge[side].alpha;
                float pml_alpha = std::max(alpha, kEpsilon);
                int xo = 0, yo = 0, xs = 0, ys = 0;
                switch (side) {
                    case kLeft: xo = -N; yo = vr.a - u.tly; xs = N; ys = vr.b - vr.a; break;
                    case kRight: xo = u.nx; yo = vr.a - u.tly; xs = N; ys = vr.b - vr.a; break;
                    case kTop: xo = vr.a - u.tlx; yo = -N; xs = vr.b - vr.a; ys = N; break;
                    case kBottom: xo = vr.a - u.tlx; yo = u.ny; xs = vr.b - vr.a; ys = N; break;
                }
                bool full = vr.a == edge.a && vr.b == edge.b;
                int ptlx = u.tlx + xo, ptly = u.tly + yo;
                int pi = make_domain(pml_alpha, ptlx, ptly, xs, ys, true, ui, nullptr);
                first_order.push_back(pi);
                if (!full) pool_[pi].local = true;
                if (alpha > 0 && full) {
                    for (int k : {(i + 1) % 4, (i + 3) % 4}) {
                        Side sd = order[k];
                        float other = pool_[ui].edge[sd].alpha;
                        float sa = std::min(std::max(kEpsilon, other), std::max(kEpsilon, pml_alpha));
                        int sx = 0, sy = 0;
                        switch (sd) {
                            case kLeft: sx = -N; break;
                            case kRight: sx = pool_[pi].nx; break;
                            case kTop: sy = -N; break;
                            case kBottom: sy = pool_[pi].ny; break;
                        }
                        int si = make_domain(sa, ptlx + sx, ptly + sy, N, N, true, pi, nullptr);
                        second_order.push_back({si, sd});
                    }
                }
            }
        }
    }
    for (int pi : first_order) add_domain(pi);
    // group corner candidates by geometry; merge twins that share a grandparent
    std::map<std::vector<int>, std::vector<int>> by_corner;
    for (auto &e : second_order) {
        int parent = pool_[e.first].pml_for[0];
        if (!pool_[parent].nb[e.second].empty()) continue;
        const DomainModel &c = pool_[e.first];
        by_corner[{c.tlx, c.tly, c.brx(), c.bry()}].push_back(e.first);
    }
    auto singular_parent = [&](int d) -> int {
        if (d >= 0 && pool_[d].is_pml && pool_[d].pml_for.size() == 1) return pool_[d].pml_for[0];
        return -1;
    };
    std::vector<int> secondary;
    for (auto &e : by_corner) {
        std::vector<int> &v = e.second;
        if (v.size() == 1) {
            secondary.push_back(v[0]);
            continue;
        }
        std::vector<int> done;
        auto is_done = [&](int x) { return std::find(done.begin(), done.end(), x) != done.end(); };
        for (size_t i = 0; i < v.size(); ++i)
            for (size_t j = i + 1; j < v.size(); ++j) {
                if (is_done(v[i]) || is_done(v[j])) continue;
                int g1 = singular_parent(singular_parent(v[i]));
                int g2 = singular_parent(singular_parent(v[j]));
                if (g1 >= 0 && g2 >= 0 && pool_[g1].id == pool_[g2].id) {
                    done.push_back(v[i]);
                    done.push_back(v[j]);
                    for (int pf : pool_[v[j]].pml_for) pool_[v[i]].pml_for.push_back(pf);
                    secondary.push_back(v[i]);
                }
            }
        for (int x : v)
            if (!is_done(x)) secondary.push_back(x);
    }
    for (int si : secondary) add_domain(si);
}

// Attenuation flags of Domain::compute_pml_matrices (kernel/core/Domain.cpp:650-686).
void SceneModel::attenuation_flags(DomainModel &d) {
    d.reversed.clear();
    int n_all = 0, n_pml = 0;
    for (int s = 0; s < 4; ++s)
        for (int n : d.nb[s]) {
            ++n_all;
            if (pool_[n].is_pml) ++n_pml;
        }
    auto pmls_at = [&](Side s) {
        int c = 0;
        for (int n : d.nb[s])
            if (pool_[n].is_pml) ++c;
        return c;
    };
    if (d.is_secondary) {
        d.corner = n_all == 2;
        d.horizontal = !d.nb[kLeft].empty() || !d.nb[kRight].empty();
        if (d.corner) {
            if (!(d.horizontal && (!d.nb[kTop].empty() || !d.nb[kBottom].empty())))
                throw PSTDSolverError("corner PML without horizontal and vertical neighbours");
            d.reversed = {!d.nb[kLeft].empty(), !d.nb[kBottom].empty()};
        } else {
            d.reversed = {!d.nb[kLeft].empty() || !d.nb[kBottom].empty()};
        }
    } else {
        d.corner = false;
        bool air_l = !d.nb[kLeft].empty() && pmls_at(kLeft) == 0;
        bool air_r = !d.nb[kRight].empty() && pmls_at(kRight) == 0;
        bool air_b = !d.nb[kBottom].empty() && pmls_at(kBottom) == 0;
        d.horizontal = air_l || air_r;
        d.reversed = {air_l || air_b};
    }
}

// Domain::find_update_directions (kernel/core/Domain.cpp:587-616).  The
// attenuation flag it consults is computed first (openPSTD reads it before
// compute_pml_matrices assigns it).
void SceneModel::update_directions(DomainModel &d) {
    int non_pml = 0;
    for (int s = 0; s < 4; ++s)
        for (int n : d.nb[s])
            if (!pool_[n].is_pml) ++non_pml;
    for (int cd = 0; cd < 2; ++cd) {
        bool su = true;
        if (non_pml == 1 && d.is_pml) {
            if (d.horizontal && cd == 0) su = true;
            else if (!d.horizontal && cd == 1) su = true;
            else if (d.local) su = false;
            else {
                for (int s = 0; s < 4; ++s) {
                    if (d.nb[s].size() == 1 && !pool_[d.nb[s][0]].is_pml) {
                        const DomainModel &u = pool_[d.nb[s][0]];
                        // openPSTD consults the neighbour's edge in the same direction s
                        su = !(u.has_edges && u.edge[s].locally_reacting);
                    }
                }
            }
        }
        d.update[cd] = su;
    }
}

// Domain::create_attenuation_array (kernel/core/Domain.cpp:743-780), in double.
void SceneModel::pml_arrays(DomainModel &d) {
    const int N = settings_.pml_cells;
    const double ampmax = settings_.ampmax, rho_air = settings_.rho, dt = dt_;
    auto profile = [&](bool ascending, std::vector<double> &fp, std::vector<double> &fv) {
        fp.assign(N, 0);
        fv.assign(N + 1, 0);
        for (int i = 0; i < N; ++i) {
            double r = (i + 0.5) / N;
            fp[i] = std::exp(-(ampmax * r * r * r * r) * dt / rho_air);
        }
        for (int i = 0; i <= N; ++i) {
            double r = (double)i / N;
            fv[i] = std::exp(-(rho_air * ampmax * r * r * r * r) * dt);
        }
        if (!ascending) {
            std::reverse(fp.begin(), fp.end());
            std::reverse(fv.begin(), fv.end());
        }
    };
    const int nx = d.nx, ny = d.ny;
    d.pml_px.assign((size_t)nx * ny, 1.f);
    d.pml_py.assign((size_t)nx * ny, 1.f);
    d.pml_vx.assign((size_t)ny * (nx + 1), 1.f);
    d.pml_vy.assign((size_t)(ny + 1) * nx, 1.f);
    auto fill_x = [&](bool asc) {
        if (nx != N) throw PSTDSolverError("X-attenuating PML width differs from PMLCells");
        std::vector<double> fp, fv;
        profile(asc, fp, fv);
        for (int y = 0; y < ny; ++y) {
            for (int x = 0; x < nx; ++x) d.pml_px[(size_t)y * nx + x] = (float)fp[x];
            for (int x = 0; x <= nx; ++x) d.pml_vx[(size_t)y * (nx + 1) + x] = (float)fv[x];
        }
    };
    auto fill_y = [&](bool asc) {
        if (ny != N) throw PSTDSolverError("Y-attenuating PML height differs from PMLCells");
        std::vector<double> fp, fv;
        profile(asc, fp, fv);
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) d.pml_py[(size_t)y * nx + x] = (float)fp[y];
        for (int y = 0; y <= ny; ++y)
            for (int x = 0; x < nx; ++x) d.pml_vy[(size_t)y * nx + x] = (float)fv[y];
    };
    if (d.is_secondary && d.corner) {
        fill_x(d.reversed.at(0));
        fill_y(d.reversed.at(1));
    } else if (d.horizontal) {
        fill_x(d.reversed.at(0));
    } else {
        fill_y(d.reversed.at(0));
    }
}

// Speaker::addDomainContribution (kernel/core/Speaker.cpp:21-39) for every
// domain, PML included (kernel/core/Scene.cpp:267-279); evaluated in double.
void SceneModel::add_speaker(float x, float y) {
    const float sx = x - 0.5f, sy = y - 0.5f;
    const double dx = settings_.grid_spacing, beta = settings_.band_width;
    for (int li : list_) {
        DomainModel &d = pool_[li];
        const double rx = (double)(sx - (float)d.tlx), ry = (double)(sy - (float)d.tly);
        for (int j = 0; j < d.ny; ++j)
            for (int i = 0; i < d.nx; ++i) {
                double ax = rx - i, ay = ry - j;
                double sq = (ax * dx) * (ax * dx) + (ay * dx) * (ay * dx);
                double p = std::exp(-beta * sq);
                double ang = std::atan2(ax, ay);
                double c = std::cos(ang), s = std::sin(ang);
                size_t k = (size_t)j * d.nx + i;
                d.p0[k] = (float)((double)d.p0[k] + p);
                d.px0[k] = (float)((double)d.px0[k] + c * c * p);
                d.py0[k] = (float)((double)d.py0[k] + s * s * p);
            }
    }
}

// Scene::add_receiver (kernel/core/Scene.cpp:254-265), Receiver::compute_with_nn
// (kernel/core/Receiver.cpp:80-83).
void SceneModel::add_receiver(float x, float y, int mode) {
    if (mode != PSTD_RECEIVER_PHYSICAL && mode != PSTD_RECEIVER_REFERENCE && mode != PSTD_RECEIVER_SPECTRAL)
        throw PSTDSolverError("unknown receiver mode " + std::to_string(mode));
    int container = -1;
    for (int li = 0; li < (int)list_.size(); ++li) {
        const DomainModel &d = pool_[list_[li]];
        if (d.is_pml) continue;
        if ((float)d.tlx <= x && x <= (float)d.brx() && (float)d.tly <= y && y <= (float)d.bry()) container = li;
    }
    if (container < 0) throw PSTDSolverError("receiver outside every domain");
    const DomainModel &d = pool_[list_[container]];
    ReceiverModel r;
    r.domain = container;
    r.cell_x = (int)x;
    r.cell_y = (int)y;
    // Receiver::compute_with_si shifts by location - grid cell (the intended
    // grid_offset; kernel/core/Receiver.cpp:41 stores the cell itself there)
    r.spectral = mode == PSTD_RECEIVER_SPECTRAL;
    r.off_x = x - (float)r.cell_x;
    r.off_y = y - (float)r.cell_y;
    int relx = r.cell_x - d.tlx, rely = r.cell_y - d.tly;
    int row = rely, col = relx;
    if (mode == PSTD_RECEIVER_REFERENCE) {
        // p0(rel.x, rel.y) of openPSTD's column-major ny x nx array: element
        // rel.x + rel.y * ny of its storage, without the bounds check of an
        // NDEBUG build (for rel.x < ny the transposed cell)
        const long long lin = (long long)relx + (long long)rely * d.ny;
        if (relx < 0 || rely < 0 || lin >= (long long)d.nx * d.ny)
            throw PSTDSolverError("receiver: openPSTD's p0(rel.x, rel.y) read lies outside its domain's pressure array");
        row = (int)(lin % d.ny);
        col = (int)(lin / d.ny);
    }
    if (row < 0 || row >= d.ny || col < 0 || col >= d.nx)
        throw PSTDSolverError("receiver cell outside its domain's pressure array");
    r.index = row * d.nx + col;
    receivers_.push_back(r);
}

// Domain::calc neighbour-pair loop (kernel/core/Domain.cpp:147-370), resolved once.
void SceneModel::build_segments() {
    segments_.clear();
    for (int li = 0; li < (int)list_.size(); ++li) {
        const DomainModel &d = pool_[list_[li]];
        if (d.rigid()) continue;
        for (int cd = 0; cd < 2; ++cd) {
            if (!d.update[cd]) continue;
            for (int ct = 0; ct < 2; ++ct) line_segments(li, cd, ct, 0);
        }
    }
    // spectral receivers read the X-shifted pressure of the neighbours of
    // their Y line (Receiver::calc_domain_fields, kernel/core/Receiver.cpp:
    // 117-121); a neighbour the step does not update along X gets its X
    // pressure lines as interpolation-only segments
    for (const ReceiverModel &r : receivers_) {
        if (!r.spectral) continue;
        const DomainModel &c = pool_[list_[r.domain]];
        const int col = r.cell_x - c.tlx;
        const pstd_plan_segment *sy = nullptr;
        for (const pstd_plan_segment &s : segments_)
            if (s.domain == r.domain && s.direction == 1 && s.field == 0 && s.line_begin <= col && col < s.line_end)
                sy = &s;
        if (!sy) throw PSTDSolverError("spectral receiver: its domain has no Y pressure line through it");
        const int sides[2] = {sy->side1, sy->side2};
        for (int nb : sides) {
            if (nb < 0) continue;
            bool has_x = false;
            for (const pstd_plan_segment &s : segments_)
                if (s.domain == nb && s.direction == 0 && s.field == 0) has_x = true;
            if (!has_x) line_segments(nb, 0, 0, 1);
        }
    }
}

// The segments of Domain::calc for (domain list index li, direction cd,
// field ct).
void SceneModel::line_segments(int li, int cd, int ct, int interp_only) {
    const DomainModel &d = pool_[list_[li]];
    const bool X = cd == 0;
    Interval own = X ? Interval{d.tly, d.bry()} : Interval{d.tlx, d.brx()};
    std::vector<char> left(own.b - own.a, 1);  // lines not yet assigned
    const std::vector<int> &n1 = X ? d.nb[kLeft] : d.nb[kBottom];
    const std::vector<int> &n2 = X ? d.nb[kRight] : d.nb[kTop];
    // openPSTD's Domain::calc sets d1 = shared_from_this() in the
    // first segment without a side-1 neighbour and never resets it
    // inside the j loop (kernel/core/Domain.cpp:148-150,210-233), so
    // a second such segment would read the domain's own field as its
    // side 1 with rho_self.  Scenes reaching that are rejected.
    int no_side1_segments = 0;
    for (size_t i = 0; i <= n1.size(); ++i) {
        int d1 = i < n1.size() ? n1[i] : -1;
        for (size_t j = 0; j <= n2.size(); ++j) {
            int d2 = j < n2.size() ? n2[j] : -1;
            auto range = [&](int pi) {
                const DomainModel &o = pool_[pi];
                return X ? Interval{o.tly, o.bry()} : Interval{o.tlx, o.brx()};
            };
            Interval lim = own;
            if (d1 >= 0) lim = intersect(lim, range(d1));
            if (d2 >= 0) lim = intersect(lim, range(d2));
            int lo = 0, hi = 0, count = 0;
            for (int v = std::max(lim.a, own.a); v < lim.b; ++v)
                if (left[v - own.a]) {
                    if (count == 0) lo = v;
                    hi = v;
                    ++count;
                }
            if (count == 0) continue;
            if (hi - lo + 1 != count)
                throw PSTDSolverError("non-contiguous segment in Domain::calc is not supported (domain " +
                                      std::to_string(d.id) + ", lines " + std::to_string(lo) + ".." +
                                      std::to_string(hi) + ", " + std::to_string(count) + " free)");
            for (int v = lo; v <= hi; ++v) left[v - own.a] = 0;
            if (d1 < 0 && ++no_side1_segments > 1)
                throw PSTDSolverError(
                    "two line segments without a side-1 neighbour in one Domain::calc pass (domain " +
                    std::to_string(d.id) + ", " + (X ? "X" : "Y") +
                    "): openPSTD reuses the domain itself as side 1 there; not supported");
            int primary = X ? d.nx : d.ny;
            int wlen = window_;
            while (wlen > primary) wlen /= 2;
            int n_total = 2 * wlen + primary + (ct == 0 ? 1 : 0);
            int table_len = next_pow2(n_total);
            int main_len = ct == 0 ? primary : primary + 1;
            int F = next_pow2(main_len + 2 * wlen);
            if (table_len < F)
                throw PSTDSolverError("domain size with n+2w a power of two: openPSTD reads its "
                                      "wavenumber table past the end");
            if (table_len != F)
                throw PSTDSolverError("domain size with n+2w a power of two: openPSTD's wavenumber "
                                      "table and FFT length disagree");
            pstd_plan_segment s{};
            s.domain = li;
            s.direction = cd;
            s.field = ct;
            s.line_begin = lo - own.a;
            s.line_end = hi + 1 - own.a;
            s.wlen = wlen;
            s.table_len = table_len;
            double r1 = d1 >= 0 ? pool_[d1].rho : d.rho;
            double r2 = d2 >= 0 ? pool_[d2].rho : d.rho;
            float c[4];
            stencil(r1, d.rho, r2, ct, c);
            s.c_side1 = c[0];
            s.c_main1 = c[1];
            s.c_side2 = c[2];
            s.c_main2 = c[3];
            s.interp_only = interp_only;
            s.side1 = d1 >= 0 ? list_index_of_pool(d1) : -1;
            s.side2 = d2 >= 0 ? list_index_of_pool(d2) : -1;
            s.side1_offset = d1 >= 0 ? (X ? d.tly - pool_[d1].tly : d.tlx - pool_[d1].tlx) : 0;
            s.side2_offset = d2 >= 0 ? (X ? d.tly - pool_[d2].tly : d.tlx - pool_[d2].tlx) : 0;
            for (int dd : {d1, d2}) {
                if (dd < 0) continue;
                int len = (X ? pool_[dd].nx : pool_[dd].ny) + (ct == 1 ? 1 : 0);
                if (len < wlen + (ct == 1 ? 1 : 0))
                    throw PSTDSolverError("neighbour narrower than the window");
            }
            segments_.push_back(s);
        }
    }
}

void SceneModel::build(OpenPSTD::Kernel::PSTDConfiguration &conf) {
    using namespace OpenPSTD::Kernel;
    PSTDSettings &st = conf.Settings;
    settings_ = pstd_settings{};
    settings_.calctime = st.GetRenderTime();
    settings_.c1 = st.GetSoundSpeed();
    settings_.ampmax = st.GetAttenuationOfPMLCells();
    settings_.freqmax = st.GetMaxFrequency();
    settings_.rho = st.GetDensityOfAir();
    settings_.patcherror = st.GetPatchError();
    settings_.tfactRK = st.GetFactRK();
    settings_.grid_spacing = st.GetGridSpacing();
    settings_.band_width = st.GetBandWidth();
    settings_.wave_length = st.GetWaveLength();
    std::vector<float> rk = st.GetRKCoefficients();
    if (rk.size() != 6) throw PSTDSolverError("the RK loop needs exactly 6 coefficients");
    for (int i = 0; i < 6; ++i) settings_.rk_coefficients[i] = rk[i];
    settings_.spectral_interpolation = st.GetSpectralInterpolation();
    settings_.pml_cells = st.GetPMLCells();
    settings_.save_nth = st.GetSaveNth();
    if (settings_.grid_spacing <= 0 || settings_.c1 <= 0) throw PSTDSolverError("grid spacing and sound speed must be positive");
    if (settings_.save_nth <= 0) throw PSTDSolverError("SaveNth must be positive");
    dt_ = st.GetTimeStep();
    window_ = st.GetWindowSize();
    num_steps_ = (int)(st.GetRenderTime() / st.GetTimeStep());
    pool_.clear();
    list_.clear();
    receivers_.clear();
    next_id_ = 0;
    const float dx = settings_.grid_spacing;
    // PSTDKernel::add_domains (kernel/PSTDKernel.cpp:73-98)
    std::vector<int> users;
    for (DomainConf &dc : conf.Domains) {
        QVector2D tl = dc.TopLeft / dx, sz = dc.Size / dx;
        // translate_edge_parameters (kernel/PSTDKernel.cpp:181-190): kernel TOP <- B, BOTTOM <- T
        EdgeParam e[4];
        e[kLeft] = {dc.L.LR, dc.L.Absorption};
        e[kRight] = {dc.R.LR, dc.R.Absorption};
        e[kTop] = {dc.B.LR, dc.B.Absorption};
        e[kBottom] = {dc.T.LR, dc.T.Absorption};
        users.push_back(make_domain(1.f, (int)tl[0], (int)tl[1], (int)sz[0], (int)sz[1], false, -1, e));
    }
    for (int u : users) add_domain(u);
    add_pml_domains();
    for (int li : list_) {
        DomainModel &d = pool_[li];
        if (d.is_pml) attenuation_flags(d);
        update_directions(d);
    }
    for (int li : list_) {
        DomainModel &d = pool_[li];
        size_t n = (size_t)d.nx * d.ny;
        d.p0.assign(n, 0.f);
        d.px0.assign(n, 0.f);
        d.py0.assign(n, 0.f);
    }
    for (QVector3D &s : conf.Speakers) {
        QVector3D g = s / dx;
        add_speaker(g[0], g[1]);
    }
    for (QVector3D &r : conf.Receivers) {
        QVector3D g = r / dx;
        add_receiver(g[0], g[1], conf.ReceiverMode);
    }
    for (int li : list_)
        if (pool_[li].is_pml) pml_arrays(pool_[li]);
    build_segments();
}

void SceneModel::plan(pstd_plan_desc *desc, std::vector<pstd_plan_domain> &doms,
                      std::vector<pstd_plan_receiver> &rcv) const {
    doms.clear();
    for (int li : list_) {
        const DomainModel &d = pool_[li];
        pstd_plan_domain p{};
        p.nx = d.nx;
        p.ny = d.ny;
        p.rho = d.rho;
        p.is_pml = d.is_pml;
        p.is_rigid = d.rigid();
        if (d.is_pml) {
            p.pml_px = d.pml_px.data();
            p.pml_py = d.pml_py.data();
            p.pml_vx = d.pml_vx.data();
            p.pml_vy = d.pml_vy.data();
        }
        doms.push_back(p);
    }
    rcv.clear();
    for (const ReceiverModel &r : receivers_) {
        pstd_plan_receiver p{};
        p.domain = r.domain;
        p.index = r.index;
        p.spectral = r.spectral ? 1 : 0;
        p.offset_x = r.off_x;
        p.offset_y = r.off_y;
        rcv.push_back(p);
    }
    *desc = pstd_plan_desc{};
    desc->n_domains = (int)doms.size();
    desc->domains = doms.data();
    desc->n_segments = (int)segments_.size();
    desc->segments = segments_.data();
    desc->n_receivers = (int)rcv.size();
    desc->receivers = rcv.data();
    desc->dx = settings_.grid_spacing;
    desc->dt = dt_;
    desc->c1 = settings_.c1;
    desc->patch_error = settings_.patcherror;
    for (int i = 0; i < 6; ++i) desc->rk_alpha[i] = settings_.rk_coefficients[i];
}

}  // namespace pstdamd
