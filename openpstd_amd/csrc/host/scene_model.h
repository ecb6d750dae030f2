// Host-side scene model of the MI355X solver: turns a PSTDConfiguration into
// the flat plan the device consumes (include/openpstd_amd.h, pstd_plan_desc).
//
// It reproduces openPSTD's scene construction (kernel/PSTDKernel.cpp:62-119,
// kernel/core/Scene.cpp:41-380, kernel/core/Domain.cpp:53-92, 543-780,
// kernel/core/Speaker.cpp:21-39, kernel/core/Receiver.cpp:34-83) and the
// neighbour/segment resolution of Domain::calc (kernel/core/Domain.cpp:95-370),
// precomputed once instead of per call.  Tables are assembled in double and
// rounded once to float.
#pragma once

#include <string>
#include <vector>

#include "../../../include/openpstd_amd.h"
#include "kernel_api.h"

namespace pstdamd {

enum Side { kLeft = 0, kRight = 1, kTop = 2, kBottom = 3 };

struct EdgeParam {
    bool locally_reacting = false;
    float alpha = 0.f;
};

struct DomainModel {
    int id = 0;
    int tlx = 0, tly = 0, nx = 0, ny = 0;  // grid cells
    float alpha = 1.f;
    float impedance = 1.f;
    float rho = 0.f;
    bool is_pml = false, is_secondary = false, local = false;
    bool has_edges = false;
    EdgeParam edge[4];
    std::vector<int> pml_for;   // pool indices
    std::vector<int> nb[4];     // pool indices per Side
    bool update[2] = {true, true};
    bool horizontal = false, corner = false;
    std::vector<bool> reversed;
    // row-major arrays: pml px/py ny*nx, vx ny*(nx+1), vy (ny+1)*nx
    std::vector<float> pml_px, pml_py, pml_vx, pml_vy;
    std::vector<float> p0, px0, py0;  // initial fields (speakers)

    int brx() const { return tlx + nx; }
    int bry() const { return tly + ny; }
    bool rigid() const { return impedance > 1000.f; }  // Domain::is_rigid (Domain.cpp:431-433)
};

struct ReceiverModel {
    int domain = -1;  // list index
    int index = -1;   // row-major p0 index
    int cell_x = 0, cell_y = 0;
    bool spectral = false;         // PSTD_RECEIVER_SPECTRAL
    float off_x = 0, off_y = 0;    // location - cell (cells)
};

class SceneModel {
public:
    // Throws OpenPSTD::Kernel::PSTDSolverError for scenes outside the supported set.
    void build(OpenPSTD::Kernel::PSTDConfiguration &conf);

    // Domains in openPSTD's domain_list order.
    int n_domains() const { return (int)list_.size(); }
    const DomainModel &domain(int list_index) const { return pool_[list_[list_index]]; }
    int list_index_of_pool(int pool_index) const;
    const std::vector<pstd_plan_segment> &segments() const { return segments_; }
    const std::vector<ReceiverModel> &receivers() const { return receivers_; }
    int num_steps() const { return num_steps_; }
    float dt() const { return dt_; }
    const pstd_settings &settings() const { return settings_; }
    // Fills `desc` with pointers into this model (valid while it lives).
    void plan(pstd_plan_desc *desc, std::vector<pstd_plan_domain> &domains_out,
              std::vector<pstd_plan_receiver> &receivers_out) const;
    std::vector<std::string> log;

private:
    int new_id() { return next_id_++; }
    int make_domain(float alpha, int tlx, int tly, int nx, int ny, bool pml, int pml_for, const EdgeParam *edges);
    void add_domain(int pi);
    void add_pml_domains();
    bool is_neighbour_of(int a, int b) const;
    bool is_pml_for(int a, int b) const;
    void attenuation_flags(DomainModel &d);
    void update_directions(DomainModel &d);
    void pml_arrays(DomainModel &d);
    void add_speaker(float x, float y);
    void add_receiver(float x, float y, int mode);
    void build_segments();
    void line_segments(int li, int cd, int ct, int interp_only);

    pstd_settings settings_{};
    float dt_ = 0.f;
    int window_ = 32;
    int num_steps_ = 0;
    int next_id_ = 0;
    std::vector<DomainModel> pool_;
    std::vector<int> list_;  // domain_list: pool indices
    std::vector<ReceiverModel> receivers_;
    std::vector<pstd_plan_segment> segments_;
};

}  // namespace pstdamd
