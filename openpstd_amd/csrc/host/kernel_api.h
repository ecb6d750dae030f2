// openPSTD's kernel API, restated without Qt for the MI355X solver.
//
// Same classes, members and method names as kernel/KernelInterface.h:15-358 and
// kernel/GeneralTypes.h:42-48, so a caller of openPSTD's kernel (its CLI
// RunCommand, CLI/main.cpp:431-449, or GUI SimulateLOperation,
// GUI/operations/long/SimulateLOperation.cpp:58-92) compiles against it
// unchanged apart from the QVector2D/QVector3D stand-ins below.
#pragma once

#include <cmath>
#include <exception>
#include <memory>
#include <string>
#include <vector>

namespace OpenPSTD {
namespace Kernel {

using PSTD_FRAME_UNIT = float;
using PSTD_FRAME = std::vector<PSTD_FRAME_UNIT>;
using PSTD_FRAME_PTR = std::shared_ptr<PSTD_FRAME>;
using PSTD_RECEIVER_DATA_UNIT = float;
using PSTD_RECEIVER_DATA = std::vector<PSTD_FRAME_UNIT>;
using PSTD_RECEIVER_DATA_PTR = std::shared_ptr<PSTD_FRAME>;

// Minimal stand-ins for Qt's QVector2D / QVector3D (float components).
class QVector2D {
public:
    QVector2D() : v{0, 0} {}
    QVector2D(float x, float y) : v{x, y} {}
    float x() const { return v[0]; }
    float y() const { return v[1]; }
    float operator[](int i) const { return v[i]; }
    friend QVector2D operator/(const QVector2D &a, float d) { return QVector2D(a.v[0] / d, a.v[1] / d); }

private:
    float v[2];
};

class QVector3D {
public:
    QVector3D() : v{0, 0, 0} {}
    QVector3D(float x, float y, float z) : v{x, y, z} {}
    float x() const { return v[0]; }
    float y() const { return v[1]; }
    float z() const { return v[2]; }
    float operator[](int i) const { return v[i]; }
    friend QVector3D operator/(const QVector3D &a, float d) { return QVector3D(a.v[0] / d, a.v[1] / d, a.v[2] / d); }

private:
    float v[3];
};

enum class CALLBACKSTATUS { STARTING, RUNNING, FINISHED };

enum PSTD_DOMAIN_SIDE {
    PSTD_DOMAIN_SIDE_TOP = 1,
    PSTD_DOMAIN_SIDE_BOTTOM = 2,
    PSTD_DOMAIN_SIDE_LEFT = 4,
    PSTD_DOMAIN_SIDE_RIGHT = 8,
    PSTD_DOMAIN_SIDE_ALL = 8 + 4 + 2 + 1,
    PSTD_DOMAIN_SIDE_NONE = 0
};

// kernel/KernelInterface.h:44-158
class PSTDSettings {
public:
    float GetGridSpacing() { return gridSpacing; }
    void SetGridSpacing(float v) { gridSpacing = v; }
    float GetPatchError() { return patcherror; }
    void SetPatchError(float v) { patcherror = v; }
    int GetWindowSize() {
        int tmp = (int)std::round((patcherror * 0.7 - 17) / 2.);
        tmp = (int)std::round(tmp * 2);
        return tmp;
    }
    float GetRenderTime() { return calctime; }
    void SetRenderTime(float v) { calctime = v; }
    int GetPMLCells() { return PMLCells; }
    void SetPMLCells(int v) { PMLCells = v; }
    float GetAttenuationOfPMLCells() { return ampMax; }
    void SetAttenuationOfPMLCells(float v) { ampMax = v; }
    float GetDensityOfAir() { return rho; }
    void SetDensityOfAir(float v) { rho = v; }
    float GetMaxFrequency() { return freqMax; }
    void SetMaxFrequency(float v) { freqMax = v; }
    float GetSoundSpeed() { return c1; }
    void SetSoundSpeed(float v) { c1 = v; }
    float GetFactRK() { return tfactRK; }
    void SetFactRK(float v) { tfactRK = v; }
    int GetSaveNth() { return SaveNth; }
    void SetSaveNth(int v) { SaveNth = v; }
    float GetBandWidth() { return band_width; }
    void SetBandWidth(float v) { band_width = v; }
    bool GetSpectralInterpolation() { return spectral_interpolation; }
    void SetSpectralInterpolation(bool v) { spectral_interpolation = v; }
    float GetWaveLength() { return wave_length; }
    void SetWaveLength(float v) { wave_length = v; }
    float GetTimeStep() { return tfactRK * gridSpacing / c1; }
    std::vector<float> GetRKCoefficients() { return rk_coefficients; }
    void SetRKCoefficients(std::vector<float> c) { rk_coefficients = c; }

private:
    float calctime = 0, c1 = 0, ampMax = 0, freqMax = 0, rho = 0, patcherror = 0, tfactRK = 0, gridSpacing = 0,
          band_width = 0, wave_length = 0;
    std::vector<float> rk_coefficients;
    bool spectral_interpolation = false;
    int PMLCells = 0;
    int SaveNth = 1;
};

// kernel/KernelInterface.h:163-176
class DomainConfEdge {
public:
    float Absorption = 0;
    bool LR = false;
};

// kernel/KernelInterface.h:181-206
class DomainConf {
public:
    QVector2D TopLeft;
    QVector2D Size;
    DomainConfEdge T, L, B, R;

    float GetAbsorption(PSTD_DOMAIN_SIDE side) {
        switch (side) {
            case PSTD_DOMAIN_SIDE_TOP: return T.Absorption;
            case PSTD_DOMAIN_SIDE_BOTTOM: return B.Absorption;
            case PSTD_DOMAIN_SIDE_LEFT: return L.Absorption;
            case PSTD_DOMAIN_SIDE_RIGHT: return R.Absorption;
            default: return 0;
        }
    }
    void SetAbsorption(PSTD_DOMAIN_SIDE sides, float a) {
        if (sides & PSTD_DOMAIN_SIDE_TOP) T.Absorption = a;
        if (sides & PSTD_DOMAIN_SIDE_BOTTOM) B.Absorption = a;
        if (sides & PSTD_DOMAIN_SIDE_LEFT) L.Absorption = a;
        if (sides & PSTD_DOMAIN_SIDE_RIGHT) R.Absorption = a;
    }
    bool GetLR(PSTD_DOMAIN_SIDE side) {
        switch (side) {
            case PSTD_DOMAIN_SIDE_TOP: return T.LR;
            case PSTD_DOMAIN_SIDE_BOTTOM: return B.LR;
            case PSTD_DOMAIN_SIDE_LEFT: return L.LR;
            case PSTD_DOMAIN_SIDE_RIGHT: return R.LR;
            default: return false;
        }
    }
    void SetLR(PSTD_DOMAIN_SIDE sides, bool lr) {
        if (sides & PSTD_DOMAIN_SIDE_TOP) T.LR = lr;
        if (sides & PSTD_DOMAIN_SIDE_BOTTOM) B.LR = lr;
        if (sides & PSTD_DOMAIN_SIDE_LEFT) L.LR = lr;
        if (sides & PSTD_DOMAIN_SIDE_RIGHT) R.LR = lr;
    }
};

// kernel/KernelInterface.h:214-241
class PSTDConfiguration {
public:
    PSTDSettings Settings;
    std::vector<QVector3D> Speakers;
    std::vector<QVector3D> Receivers;
    std::vector<DomainConf> Domains;
    // Receiver indexing rule (PSTD_RECEIVER_* of openpstd_amd.h); openPSTD has none.
    int ReceiverMode = 0;

    static std::shared_ptr<PSTDConfiguration> CreateDefaultConf();
    static std::shared_ptr<PSTDConfiguration> CreateEmptyConf();
};

// kernel/KernelInterface.h:243-271
class KernelCallbackLog {
public:
    virtual ~KernelCallbackLog() = default;
    virtual void Fatal(std::string message) {}
    virtual void Error(std::string message) {}
    virtual void Warning(std::string message) {}
    virtual void Info(std::string message) {}
    virtual void Debug(std::string message) {}
};

// kernel/KernelInterface.h:279-304
class KernelCallback : public KernelCallbackLog {
public:
    virtual void Callback(CALLBACKSTATUS status, std::string message, int frame) {}
    virtual void WriteFrame(int frame, int domain, PSTD_FRAME_PTR data) = 0;
    virtual void WriteSample(int startSample, int receiver, std::vector<float> data) = 0;
};

// kernel/KernelInterface.h:309-327
class SimulationMetadata {
public:
    std::vector<std::vector<int>> DomainMetadata;
    int Framecount = 0;
    std::vector<std::vector<int>> DomainPositions;
};

// kernel/KernelInterface.h:332-351
class KernelInterface {
public:
    virtual ~KernelInterface() = default;
    virtual void initialize_kernel(std::shared_ptr<PSTDConfiguration> config,
                                   std::shared_ptr<KernelCallbackLog> callbackLog) = 0;
    virtual void run(std::shared_ptr<KernelCallback> callback) = 0;
    virtual SimulationMetadata get_metadata() = 0;
};

// kernel/KernelInterface.h:353-356
class PSTDKernelNotConfiguredException : public std::exception {
public:
    const char *what() const noexcept override { return "Kernel is not yet configured"; }
};

// Raised by the GPU solver for scenes it cannot run (and for run() with GPU=false).
class PSTDSolverError : public std::exception {
public:
    explicit PSTDSolverError(std::string m, int code = -1) : msg(std::move(m)), status(code) {}
    const char *what() const noexcept override { return msg.c_str(); }
    std::string msg;
    int status;
};

}  // namespace Kernel
}  // namespace OpenPSTD
