// PSTDKernel for the MI355X solver: openPSTD's kernel facade (kernel/PSTDKernel.h:52-150)
// with the GPU solver slot filled (kernel/Solver.cpp:153-160).
#pragma once

#include <memory>

#include "kernel_api.h"
#include "scene_model.h"

namespace OpenPSTD {
namespace Kernel {

// The device solver: the body openPSTD's GPUSingleThreadSolver::compute_propagation
// lacks.  Same output contract as SingleThreadSolver::compute_propagation
// (kernel/Solver.cpp:62-93): Info messages, WriteFrame for non-PML domains and
// WriteSample per receiver when frame % SaveNth == 0.
class GPUSolver {
public:
    GPUSolver(const pstdamd::SceneModel &scene, std::shared_ptr<KernelCallback> callback, int device = 0,
              int form = PSTD_FORM_FFT);
    ~GPUSolver();
    GPUSolver(const GPUSolver &) = delete;
    GPUSolver &operator=(const GPUSolver &) = delete;
    void compute_propagation();
    // creates the device context, uploads the plan and the initial fields
    void prepare();
    // device context after compute_propagation (state of the last frame)
    pstd_ctx *context() const { return ctx; }

private:
    const pstdamd::SceneModel &scene;
    std::shared_ptr<KernelCallback> callback;
    int device;
    int form;
    pstd_ctx *ctx = nullptr;
};

class PSTDKernel : public KernelInterface {
public:
    // kernel/PSTDKernel.cpp:44-47.  GPU selects the MI355X solver (MCPU is
    // accepted and ignored, as GPUMultiThreadSolver is the same device path).
    PSTDKernel(bool GPU, bool MCPU) : GPU(GPU), MCPU(MCPU) {}
    void initialize_kernel(std::shared_ptr<PSTDConfiguration> config,
                           std::shared_ptr<KernelCallbackLog> callbackLog) override;
    void run(std::shared_ptr<KernelCallback> callback) override;
    SimulationMetadata get_metadata() override;
    const pstdamd::SceneModel &get_scene() const { return scene; }
    bool configured() const { return (bool)config; }
    bool gpu() const { return GPU; }
    // derivative form of the solvers created from now on (PSTD_FORM_*)
    void set_form(int f) { form = f; }
    int get_form() const { return form; }
    // solver of the last run (its device state stays readable until the next run)
    const GPUSolver *last_solver() const { return solver.get(); }
    // a solver with its device state initialised (frame 0 next), not run
    GPUSolver *prepare_solver();

private:
    bool GPU, MCPU;
    int form = PSTD_FORM_FFT;
    std::shared_ptr<PSTDConfiguration> config;
    std::shared_ptr<KernelCallbackLog> callbackLog;
    pstdamd::SceneModel scene;
    std::unique_ptr<GPUSolver> solver;
};

}  // namespace Kernel
}  // namespace OpenPSTD
