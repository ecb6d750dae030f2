// Test driver: openPSTD's C++ kernel API as its CLI uses it (CLI/main.cpp:431-449):
// PSTDConfiguration -> PSTDKernel(GPU=true, MCPU=false) -> initialize_kernel ->
// run(callback) -> get_metadata, compiled against the restated header
// openpstd_amd/csrc/host/kernel_api.h and linked to libopenpstd_amd.so only.
//
//   kernel_api_run <config.txt> <out.txt>
//
// config.txt (one item per line, floats written to round-trip float32):
//   settings calctime c1 ampmax freqmax rho patcherror tfactRK dx bandwidth wavelength pmlcells saventh
//   rk a0 a1 a2 a3 a4 a5
//   domain tlx tly sx sy T L B R
//   speaker x y z
//   receiver x y z
// out.txt: "frame f d n v..." per WriteFrame, "sample s r v" per WriteSample,
// "log level message" per log call, "meta n_domains framecount", "domain nx ny nz".
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>

#include "pstd_kernel.h"

using namespace OpenPSTD::Kernel;

namespace {

class Recorder : public KernelCallback {
public:
    explicit Recorder(std::ostream &o) : out(o) {}
    void WriteFrame(int frame, int domain, PSTD_FRAME_PTR data) override {
        out << "frame " << frame << " " << domain << " " << data->size();
        char buf[32];
        for (float v : *data) {
            std::snprintf(buf, sizeof buf, " %.9g", v);
            out << buf;
        }
        out << "\n";
    }
    void WriteSample(int startSample, int receiver, std::vector<float> data) override {
        char buf[32];
        std::snprintf(buf, sizeof buf, "%.9g", data.at(0));
        out << "sample " << startSample << " " << receiver << " " << buf << "\n";
    }
    void Info(std::string message) override { out << "log info " << message << "\n"; }
    void Error(std::string message) override { out << "log error " << message << "\n"; }

private:
    std::ostream &out;
};

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3) {
        std::cerr << "usage: kernel_api_run <config.txt> <out.txt>\n";
        return 2;
    }
    auto conf = PSTDConfiguration::CreateEmptyConf();
    std::ifstream in(argv[1]);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream s(line);
        std::string key;
        s >> key;
        if (key == "settings") {
            float calctime, c1, ampmax, freqmax, rho, pe, tfact, dx, bw, wl;
            int pml, nth;
            s >> calctime >> c1 >> ampmax >> freqmax >> rho >> pe >> tfact >> dx >> bw >> wl >> pml >> nth;
            PSTDSettings &st = conf->Settings;
            st.SetRenderTime(calctime);
            st.SetSoundSpeed(c1);
            st.SetAttenuationOfPMLCells(ampmax);
            st.SetMaxFrequency(freqmax);
            st.SetDensityOfAir(rho);
            st.SetPatchError(pe);
            st.SetFactRK(tfact);
            st.SetGridSpacing(dx);
            st.SetBandWidth(bw);
            st.SetWaveLength(wl);
            st.SetPMLCells(pml);
            st.SetSaveNth(nth);
            st.SetSpectralInterpolation(false);
        } else if (key == "rk") {
            std::vector<float> a(6);
            for (float &v : a) s >> v;
            conf->Settings.SetRKCoefficients(a);
        } else if (key == "domain") {
            float x, y, w, h, T, L, B, R;
            s >> x >> y >> w >> h >> T >> L >> B >> R;
            DomainConf d;
            d.TopLeft = QVector2D(x, y);
            d.Size = QVector2D(w, h);
            d.T.Absorption = T;
            d.L.Absorption = L;
            d.B.Absorption = B;
            d.R.Absorption = R;
            conf->Domains.push_back(d);
        } else if (key == "speaker" || key == "receiver") {
            float x, y, z;
            s >> x >> y >> z;
            (key == "speaker" ? conf->Speakers : conf->Receivers).push_back(QVector3D(x, y, z));
        }
    }
    std::ofstream out(argv[2]);
    try {
        PSTDKernel kernel(true, false);
        bool threw = false;
        try {
            kernel.get_metadata();
        } catch (const PSTDKernelNotConfiguredException &) {
            threw = true;  // before initialize_kernel, as openPSTD's PSTDKernel
        }
        out << "not_configured_throws " << (threw ? 1 : 0) << "\n";
        kernel.initialize_kernel(conf, std::make_shared<KernelCallbackLog>());
        auto rec = std::make_shared<Recorder>(out);
        kernel.run(rec);
        SimulationMetadata meta = kernel.get_metadata();
        out << "meta " << meta.DomainMetadata.size() << " " << meta.Framecount << "\n";
        for (auto &d : meta.DomainMetadata) out << "domain " << d.at(0) << " " << d.at(1) << " " << d.at(2) << "\n";
    } catch (const std::exception &e) {
        out << "exception " << e.what() << "\n";
        std::cerr << "exception: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
