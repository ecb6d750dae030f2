// PSTDKernel and the GPU solver of the MI355X build (see pstd_kernel.h).
#include "pstd_kernel.h"

#include <algorithm>
#include <string>
#include <vector>

namespace OpenPSTD {
namespace Kernel {

// kernel/KernelInterface.cpp:196-249
std::shared_ptr<PSTDConfiguration> PSTDConfiguration::CreateDefaultConf() {
    auto conf = std::make_shared<PSTDConfiguration>();
    conf->Settings.SetRenderTime(1.0f);
    float sound_speed = 340;
    float grid_spacing = 0.2f;
    conf->Settings.SetSoundSpeed(sound_speed);
    conf->Settings.SetAttenuationOfPMLCells(20000.0f);
    conf->Settings.SetDensityOfAir(1.2f);
    conf->Settings.SetPatchError(70.0f);
    conf->Settings.SetFactRK(0.5f);
    conf->Settings.SetGridSpacing(grid_spacing);
    conf->Settings.SetPMLCells(50);
    conf->Settings.SetSaveNth(1);
    conf->Settings.SetBandWidth((float)(3 * 1e-6 * (sound_speed / grid_spacing) * (sound_speed / grid_spacing)));
    conf->Settings.SetRKCoefficients({1.179799016570605e-1f, 1.846469664911166e-1f, 2.466236043095944e-1f,
                                      3.318395427360000e-1f, 5e-1f, 1.f});
    conf->Settings.SetSpectralInterpolation(true);
    conf->Speakers.push_back(QVector3D(4, 5, 0));
    conf->Receivers.push_back(QVector3D(6, 5, 0));
    DomainConf d;
    d.TopLeft = QVector2D(0, 0);
    d.Size = QVector2D(10, 15);
    d.T.Absorption = 1;
    d.B.Absorption = 0;
    d.L.Absorption = .3f;
    d.R.Absorption = 1;
    DomainConf d2;
    d2.TopLeft = QVector2D(10, 4);
    d2.Size = QVector2D(20, 15);
    d2.T.Absorption = 1;
    d2.B.Absorption = 0;
    d2.L.Absorption = 1;
    d2.R.Absorption = 1;
    conf->Domains.push_back(d);
    conf->Domains.push_back(d2);
    return conf;
}

// kernel/KernelInterface.cpp:251-269
std::shared_ptr<PSTDConfiguration> PSTDConfiguration::CreateEmptyConf() {
    auto conf = std::make_shared<PSTDConfiguration>();
    conf->Settings.SetRenderTime(0);
    conf->Settings.SetSoundSpeed(0);
    conf->Settings.SetAttenuationOfPMLCells(0);
    conf->Settings.SetDensityOfAir(0);
    conf->Settings.SetPatchError(0);
    conf->Settings.SetFactRK(0);
    conf->Settings.SetGridSpacing(0);
    conf->Settings.SetPMLCells(0);
    conf->Settings.SetSaveNth(0);
    conf->Settings.SetBandWidth(0);
    conf->Settings.SetSpectralInterpolation(false);
    return conf;
}

GPUSolver::GPUSolver(const pstdamd::SceneModel &scene, std::shared_ptr<KernelCallback> callback, int device, int form)
    : scene(scene), callback(std::move(callback)), device(device), form(form) {}

GPUSolver::~GPUSolver() {
    if (ctx) pstd_ctx_destroy(ctx);
}

static void check_rc(pstd_ctx *ctx, int rc, const char *what) {
    if (rc != PSTD_OK)
        throw PSTDSolverError(std::string(what) + ": " + (ctx ? pstd_ctx_last_error(ctx) : "no HIP device"), rc);
}

void GPUSolver::prepare() {
    auto check = [&](int rc, const char *what) { check_rc(ctx, rc, what); };
    if (!ctx) check(pstd_ctx_create(device, &ctx), "pstd_ctx_create");
    check(pstd_ctx_set_form(ctx, form), "pstd_ctx_set_form");
    pstd_plan_desc desc;
    std::vector<pstd_plan_domain> doms;
    std::vector<pstd_plan_receiver> rcv;
    scene.plan(&desc, doms, rcv);
    check(pstd_ctx_upload_plan(ctx, &desc), "pstd_ctx_upload_plan");
    for (int i = 0; i < scene.n_domains(); ++i) {
        const pstdamd::DomainModel &d = scene.domain(i);
        check(pstd_ctx_set_field(ctx, i, PSTD_FIELD_P0, d.p0.data(), d.p0.size()), "set p0");
        check(pstd_ctx_set_field(ctx, i, PSTD_FIELD_PX, d.px0.data(), d.px0.size()), "set px0");
        check(pstd_ctx_set_field(ctx, i, PSTD_FIELD_PY, d.py0.data(), d.py0.size()), "set py0");
    }
}

void GPUSolver::compute_propagation() {
    auto check = [&](int rc, const char *what) { check_rc(ctx, rc, what); };
    callback->Info("Starting simulation");
    prepare();
    const int nrec = (int)scene.receivers().size();
    const int steps = scene.num_steps();
    const int save_nth = scene.settings().save_nth;
    std::vector<float> samples;
    int done = 0;
    while (done < steps) {
        // advance up to the next frame that is written (frame % SaveNth == 0)
        int next_saved = ((done + save_nth - 1) / save_nth) * save_nth;
        int target = std::min(next_saved, steps - 1);
        int n = target - done + 1;
        samples.assign((size_t)n * std::max(1, nrec), 0.f);
        check(pstd_ctx_advance(ctx, n, nrec ? samples.data() : nullptr), "pstd_ctx_advance");
        for (int f = done; f < target; ++f) callback->Info("Finished frame: " + std::to_string(f));
        if (target % save_nth == 0) {
            // kernel/Solver.cpp:80-84: non-PML domains in domain_list order
            for (int i = 0; i < scene.n_domains(); ++i) {
                const pstdamd::DomainModel &d = scene.domain(i);
                if (d.is_pml) continue;
                auto frame = std::make_shared<PSTD_FRAME>((size_t)d.nx * d.ny);
                check(pstd_ctx_read_frame(ctx, i, frame->data(), frame->size()), "pstd_ctx_read_frame");
                callback->WriteFrame(target, d.id, frame);
            }
            // kernel/Solver.cpp:86-91: last received value per receiver
            for (int r = 0; r < nrec; ++r)
                callback->WriteSample(target, r, std::vector<float>{samples[(size_t)(n - 1) * nrec + r]});
        }
        callback->Info("Finished frame: " + std::to_string(target));
        done = target + 1;
    }
    callback->Info("Succesfully finished simulation");
}

// kernel/PSTDKernel.cpp:49-59
void PSTDKernel::initialize_kernel(std::shared_ptr<PSTDConfiguration> cfg, std::shared_ptr<KernelCallbackLog> log) {
    callbackLog = log ? log : std::make_shared<KernelCallbackLog>();
    callbackLog->Debug("Initializing kernel");
    solver.reset();
    config.reset();
    scene.build(*cfg);
    config = cfg;
    for (int i = 0; i < scene.n_domains(); ++i) {
        const pstdamd::DomainModel &d = scene.domain(i);
        callbackLog->Debug("Domain " + std::to_string(d.id) + (d.is_secondary ? " (sec_pml)" : d.is_pml ? " (pml)" : "") +
                           ", top left (" + std::to_string(d.tlx) + "," + std::to_string(d.tly) + "), size (" +
                           std::to_string(d.nx) + "," + std::to_string(d.ny) + ")");
    }
    callbackLog->Debug("Finished initializing kernel");
}

// kernel/PSTDKernel.cpp:121-148
void PSTDKernel::run(std::shared_ptr<KernelCallback> callback) {
    if (!config) throw PSTDKernelNotConfiguredException();
    if (!GPU)
        throw PSTDSolverError("this library provides the GPU solver only; construct PSTDKernel(GPU=true, ...)",
                              PSTD_ERR_UNSUPPORTED);
    solver.reset(new GPUSolver(scene, callback, 0, form));
    solver->compute_propagation();
}

namespace {
class NullCallback : public KernelCallback {
public:
    void WriteFrame(int, int, PSTD_FRAME_PTR) override {}
    void WriteSample(int, int, std::vector<float>) override {}
};
}  // namespace

GPUSolver *PSTDKernel::prepare_solver() {
    if (!config) throw PSTDKernelNotConfiguredException();
    if (!GPU)
        throw PSTDSolverError("this library provides the GPU solver only; construct PSTDKernel(GPU=true, ...)",
                              PSTD_ERR_UNSUPPORTED);
    solver.reset(new GPUSolver(scene, std::make_shared<NullCallback>(), 0, form));
    solver->prepare();
    return solver.get();
}

// kernel/PSTDKernel.cpp:154-168
SimulationMetadata PSTDKernel::get_metadata() {
    if (!config) throw PSTDKernelNotConfiguredException();
    SimulationMetadata result;
    for (int i = 0; i < scene.n_domains(); ++i) {
        const pstdamd::DomainModel &d = scene.domain(i);
        result.DomainMetadata.push_back({d.nx, d.ny, 0});
    }
    result.Framecount = scene.num_steps();
    return result;
}

}  // namespace Kernel
}  // namespace OpenPSTD
