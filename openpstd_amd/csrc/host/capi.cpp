// C ABI of the kernel facade (pstd_kernel_* in include/openpstd_amd.h).
#include <cstring>
#include <exception>
#include <functional>
#include <string>
#include <vector>

#include "plan_tables.h"
#include "pstd_file.h"
#include "pstd_kernel.h"

using namespace OpenPSTD::Kernel;

struct pstd_kernel {
    std::unique_ptr<PSTDKernel> k;
    std::string err;
};

struct pstd_file {
    pstdamd::PSTDFileData data;
    std::string err;
    // arrays the last pstd_file_get_scene handed out
    std::vector<float> speakers, receivers;
    std::vector<pstd_domain_conf> domains;
    // the keys in byte order, as of the last pstd_file_record_count
    std::vector<std::string> keys;
};

namespace {

class CLog : public KernelCallback {
public:
    explicit CLog(const pstd_callbacks *c) {
        if (c) cb = *c;
    }
    void Fatal(std::string m) override { log(0, m); }
    void Error(std::string m) override { log(1, m); }
    void Warning(std::string m) override { log(2, m); }
    void Info(std::string m) override { log(3, m); }
    void Debug(std::string m) override { log(4, m); }
    void WriteFrame(int frame, int domain, PSTD_FRAME_PTR data) override {
        if (cb.write_frame) cb.write_frame(cb.user, frame, domain, data->data(), data->size());
    }
    void WriteSample(int start, int receiver, std::vector<float> data) override {
        if (cb.write_sample) cb.write_sample(cb.user, start, receiver, data.data(), data.size());
    }

private:
    void log(int level, const std::string &m) {
        if (cb.log) cb.log(cb.user, level, m.c_str());
    }
    pstd_callbacks cb{};
};

std::shared_ptr<PSTDConfiguration> to_conf(const pstd_config *c) {
    auto conf = std::make_shared<PSTDConfiguration>();
    const pstd_settings &s = c->settings;
    PSTDSettings &st = conf->Settings;
    st.SetRenderTime(s.calctime);
    st.SetSoundSpeed(s.c1);
    st.SetAttenuationOfPMLCells(s.ampmax);
    st.SetMaxFrequency(s.freqmax);
    st.SetDensityOfAir(s.rho);
    st.SetPatchError(s.patcherror);
    st.SetFactRK(s.tfactRK);
    st.SetGridSpacing(s.grid_spacing);
    st.SetBandWidth(s.band_width);
    st.SetWaveLength(s.wave_length);
    st.SetRKCoefficients(std::vector<float>(s.rk_coefficients, s.rk_coefficients + 6));
    st.SetSpectralInterpolation(s.spectral_interpolation != 0);
    st.SetPMLCells(s.pml_cells);
    st.SetSaveNth(s.save_nth);
    for (int i = 0; i < c->n_speakers; ++i)
        conf->Speakers.emplace_back(c->speakers[3 * i], c->speakers[3 * i + 1], c->speakers[3 * i + 2]);
    for (int i = 0; i < c->n_receivers; ++i)
        conf->Receivers.emplace_back(c->receivers[3 * i], c->receivers[3 * i + 1], c->receivers[3 * i + 2]);
    for (int i = 0; i < c->n_domains; ++i) {
        const pstd_domain_conf &dc = c->domains[i];
        DomainConf d;
        d.TopLeft = QVector2D(dc.top_left[0], dc.top_left[1]);
        d.Size = QVector2D(dc.size[0], dc.size[1]);
        d.T = {dc.T.absorption, dc.T.lr != 0};
        d.L = {dc.L.absorption, dc.L.lr != 0};
        d.B = {dc.B.absorption, dc.B.lr != 0};
        d.R = {dc.R.absorption, dc.R.lr != 0};
        conf->Domains.push_back(d);
    }
    conf->ReceiverMode = c->receiver_mode;
    return conf;
}

template <class F>
int guarded(pstd_kernel *k, F &&f) {
    if (!k) return PSTD_ERR_INVALID;
    try {
        k->err.clear();
        return f();
    } catch (const PSTDKernelNotConfiguredException &e) {
        k->err = e.what();
        return PSTD_ERR_NOT_CONFIGURED;
    } catch (const PSTDSolverError &e) {
        k->err = e.what();
        return e.status < 0 ? e.status : PSTD_ERR_INVALID;
    } catch (const std::bad_alloc &) {
        k->err = "out of memory";
        return PSTD_ERR_ALLOC;
    } catch (const std::exception &e) {
        k->err = e.what();
        return PSTD_ERR_INVALID;
    }
}

}  // namespace

extern "C" {

int pstd_kernel_create(int gpu, int mcpu, pstd_kernel **out) {
    if (!out) return PSTD_ERR_INVALID;
    *out = new pstd_kernel();
    (*out)->k.reset(new PSTDKernel(gpu != 0, mcpu != 0));
    return PSTD_OK;
}

void pstd_kernel_destroy(pstd_kernel *k) { delete k; }

const char *pstd_kernel_last_error(const pstd_kernel *k) { return k ? k->err.c_str() : "null kernel"; }

int pstd_kernel_initialize(pstd_kernel *k, const pstd_config *conf, const pstd_callbacks *log) {
    return guarded(k, [&] {
        if (!conf) return PSTD_ERR_INVALID;
        k->k->initialize_kernel(to_conf(conf), std::make_shared<CLog>(log));
        return PSTD_OK;
    });
}

int pstd_kernel_run(pstd_kernel *k, const pstd_callbacks *cb) {
    return guarded(k, [&] {
        k->k->run(std::make_shared<CLog>(cb));
        return PSTD_OK;
    });
}

int pstd_kernel_get_metadata(pstd_kernel *k, int32_t *n_domains, int32_t *sizes_xyz, int32_t max_domains,
                             int32_t *framecount) {
    return guarded(k, [&] {
        SimulationMetadata m = k->k->get_metadata();
        if (n_domains) *n_domains = (int32_t)m.DomainMetadata.size();
        if (framecount) *framecount = m.Framecount;
        if (sizes_xyz)
            for (int i = 0; i < (int)m.DomainMetadata.size() && i < max_domains; ++i)
                for (int j = 0; j < 3; ++j) sizes_xyz[3 * i + j] = m.DomainMetadata[i][j];
        return PSTD_OK;
    });
}

int pstd_kernel_domain_info(pstd_kernel *k, int index, int32_t *info, float *finfo) {
    return guarded(k, [&] {
        if (!k->k->configured()) throw PSTDKernelNotConfiguredException();
        const pstdamd::SceneModel &s = k->k->get_scene();
        if (index < 0 || index >= s.n_domains()) return PSTD_ERR_INVALID;
        const pstdamd::DomainModel &d = s.domain(index);
        int32_t v[16] = {d.id, d.tlx, d.tly, d.nx, d.ny, d.is_pml, d.is_secondary, d.rigid(), d.local,
                         (int)d.nb[pstdamd::kLeft].size(), (int)d.nb[pstdamd::kRight].size(),
                         (int)d.nb[pstdamd::kTop].size(), (int)d.nb[pstdamd::kBottom].size(), d.update[0],
                         d.update[1], (int)d.pml_for.size()};
        if (info) std::memcpy(info, v, sizeof v);
        if (finfo) {
            finfo[0] = d.alpha;
            finfo[1] = d.impedance;
            finfo[2] = d.rho;
        }
        return PSTD_OK;
    });
}

int pstd_kernel_read_field(pstd_kernel *k, int index, int field, float *out, size_t count) {
    return guarded(k, [&] {
        if (!k->k->configured()) throw PSTDKernelNotConfiguredException();
        const pstdamd::SceneModel &s = k->k->get_scene();
        if (index < 0 || index >= s.n_domains() || !out) return PSTD_ERR_INVALID;
        const pstdamd::DomainModel &d = s.domain(index);
        size_t n = field == PSTD_FIELD_VX ? (size_t)d.ny * (d.nx + 1)
                   : field == PSTD_FIELD_VY ? (size_t)(d.ny + 1) * d.nx
                                            : (size_t)d.nx * d.ny;
        if (count != n) {
            k->err = "field size mismatch";
            return PSTD_ERR_INVALID;
        }
        const GPUSolver *sv = k->k->last_solver();
        if (sv && sv->context()) {
            int rc = pstd_ctx_get_field(sv->context(), index, field, out, count);
            if (rc) k->err = pstd_ctx_last_error(sv->context());
            return rc;
        }
        switch (field) {
            case PSTD_FIELD_P0: std::memcpy(out, d.p0.data(), n * sizeof(float)); break;
            case PSTD_FIELD_PX: std::memcpy(out, d.px0.data(), n * sizeof(float)); break;
            case PSTD_FIELD_PY: std::memcpy(out, d.py0.data(), n * sizeof(float)); break;
            case PSTD_FIELD_VX:
            case PSTD_FIELD_VY: std::memset(out, 0, n * sizeof(float)); break;
            default: return PSTD_ERR_INVALID;
        }
        return PSTD_OK;
    });
}

int pstd_kernel_read_pml(pstd_kernel *k, int index, int which, float *out, size_t count) {
    return guarded(k, [&] {
        if (!k->k->configured()) throw PSTDKernelNotConfiguredException();
        const pstdamd::SceneModel &s = k->k->get_scene();
        if (index < 0 || index >= s.n_domains() || !out || which < 0 || which > 3) return PSTD_ERR_INVALID;
        const pstdamd::DomainModel &d = s.domain(index);
        const std::vector<float> *v = which == 0 ? &d.pml_px : which == 1 ? &d.pml_py : which == 2 ? &d.pml_vx : &d.pml_vy;
        if (v->size() != count) {
            k->err = "PML array size mismatch (or not a PML domain)";
            return PSTD_ERR_INVALID;
        }
        std::memcpy(out, v->data(), count * sizeof(float));
        return PSTD_OK;
    });
}

int pstd_kernel_prepare(pstd_kernel *k, pstd_ctx **ctx) {
    return guarded(k, [&] {
        if (!ctx) return PSTD_ERR_INVALID;
        *ctx = k->k->prepare_solver()->context();
        return PSTD_OK;
    });
}

int pstd_kernel_set_form(pstd_kernel *k, int form) {
    if (!k || (form != PSTD_FORM_FFT && form != PSTD_FORM_GEMM_F16 && form != PSTD_FORM_GEMM_BF16 &&
               form != PSTD_FORM_AUTO))
        return PSTD_ERR_INVALID;
    k->k->set_form(form);
    return PSTD_OK;
}

int pstd_batch_prepare(pstd_kernel *const *kernels, int n, int device, pstd_ctx **out) {
    if (!kernels || n <= 0 || !out) return PSTD_ERR_INVALID;
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!kernels[i]) return PSTD_ERR_INVALID;
        if (!kernels[i]->k->configured()) return PSTD_ERR_NOT_CONFIGURED;
        if (!kernels[i]->k->gpu()) return PSTD_ERR_UNSUPPORTED;
    }
    std::vector<pstd_plan_desc> plans(n);
    std::vector<std::vector<pstd_plan_domain>> doms(n);
    std::vector<std::vector<pstd_plan_receiver>> rcv(n);
    for (int i = 0; i < n; ++i) kernels[i]->k->get_scene().plan(&plans[i], doms[i], rcv[i]);
    pstd_ctx *ctx = nullptr;
    int rc = pstd_ctx_create(device, &ctx);
    if (rc) return rc;
    rc = pstd_ctx_set_form(ctx, kernels[0]->k->get_form());
    if (!rc) rc = pstd_ctx_upload_plans(ctx, plans.data(), n);
    int base = 0;
    for (int i = 0; i < n && rc == PSTD_OK; ++i) {
        const pstdamd::SceneModel &s = kernels[i]->k->get_scene();
        for (int d = 0; d < s.n_domains() && rc == PSTD_OK; ++d) {
            const pstdamd::DomainModel &m = s.domain(d);
            rc = pstd_ctx_set_field(ctx, base + d, PSTD_FIELD_P0, m.p0.data(), m.p0.size());
            if (!rc) rc = pstd_ctx_set_field(ctx, base + d, PSTD_FIELD_PX, m.px0.data(), m.px0.size());
            if (!rc) rc = pstd_ctx_set_field(ctx, base + d, PSTD_FIELD_PY, m.py0.data(), m.py0.size());
        }
        base += s.n_domains();
    }
    if (rc) {
        // the message goes with the first kernel (the context is gone)
        kernels[0]->err = std::string("pstd_batch_prepare: ") + pstd_ctx_last_error(ctx);
        pstd_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return PSTD_OK;
}

long long pstd_derivative_matrix(int F, int table_len, int field, double dx, int wlen, int n_main, double patch_error,
                                 const double coef[4], double *out, size_t cap) {
    if (!coef || (field != 0 && field != 1)) return PSTD_ERR_INVALID;
    pstdamd::SegmentShape s{F, table_len, field, dx, wlen, n_main, patch_error, coef[0], coef[1], coef[2], coef[3]};
    if (!s.valid()) return PSTD_ERR_INVALID;
    try {
        const size_t n = (size_t)s.n_out() * s.K();
        if (!out) return (long long)n;
        if (cap < n) return PSTD_ERR_INVALID;
        std::vector<double> D = pstdamd::derivative_matrix(s);
        std::memcpy(out, D.data(), n * sizeof(double));
        return (long long)n;
    } catch (const std::exception &) {
        return PSTD_ERR_INVALID;
    }
}

long long pstd_shift_operator_row(int F, int table_len, double dx, int wlen, int n_main, double patch_error,
                                  const double coef[4], double offset, int row, double *out, size_t cap) {
    if (!coef) return PSTD_ERR_INVALID;
    pstdamd::SegmentShape s{F, table_len, 0, dx, wlen, n_main, patch_error, coef[0], coef[1], coef[2], coef[3]};
    if (!s.valid() || row < 0 || row >= s.n_out()) return PSTD_ERR_INVALID;
    try {
        const size_t n = (size_t)s.K();
        if (!out) return (long long)n;
        if (cap < n) return PSTD_ERR_INVALID;
        std::vector<double> r = pstdamd::operator_row(s, pstdamd::shift_filter(F, table_len, offset), row);
        std::memcpy(out, r.data(), n * sizeof(double));
        return (long long)n;
    } catch (const std::exception &) {
        return PSTD_ERR_INVALID;
    }
}

int pstd_kernel_segment_count(pstd_kernel *k) {
    if (!k || !k->k->configured()) return -1;
    return (int)k->k->get_scene().segments().size();
}

int pstd_kernel_get_segment(pstd_kernel *k, int i, pstd_plan_segment *out) {
    if (!k || !k->k->configured() || !out) return PSTD_ERR_INVALID;
    const auto &segs = k->k->get_scene().segments();
    if (i < 0 || i >= (int)segs.size()) return PSTD_ERR_INVALID;
    *out = segs[i];
    return PSTD_OK;
}

// ------------------------------------------------------------ .pstd files
static int file_guarded(pstd_file *f, const std::function<int()> &fn) {
    if (!f) return PSTD_ERR_INVALID;
    try {
        f->err.clear();
        return fn();
    } catch (const PSTDSolverError &e) {
        f->err = e.what();
        return e.status < 0 ? e.status : PSTD_ERR_INVALID;
    } catch (const std::exception &e) {
        f->err = e.what();
        return PSTD_ERR_INVALID;
    }
}

static int file_make(const char *path, bool create, pstd_file **out) {
    if (!path || !out) return PSTD_ERR_INVALID;
    *out = new pstd_file();
    return file_guarded(*out, [&] {
        if (create) (*out)->data.create(path);
        else (*out)->data.open(path);
        return PSTD_OK;
    });
}

int pstd_file_open(const char *path, pstd_file **out) { return file_make(path, false, out); }

int pstd_file_new(const char *path, pstd_file **out) { return file_make(path, true, out); }

const char *pstd_file_last_error(const pstd_file *f) { return f ? f->err.c_str() : "null file"; }

void pstd_file_close(pstd_file *f) { delete f; }

int pstd_file_get_scene(pstd_file *f, int results, pstd_config *out) {
    return file_guarded(f, [&] {
        if (!out) return PSTD_ERR_INVALID;
        const pstdamd::SceneConf c =
            f->data.scene(results ? pstdamd::PSTDFileData::kResultsScene : pstdamd::PSTDFileData::kScene);
        f->speakers.clear();
        f->receivers.clear();
        for (const auto &p : c.speakers) f->speakers.insert(f->speakers.end(), p.begin(), p.end());
        for (const auto &p : c.receivers) f->receivers.insert(f->receivers.end(), p.begin(), p.end());
        f->domains = c.domains;
        *out = pstd_config{};
        out->settings = c.settings;
        out->n_speakers = (int32_t)c.speakers.size();
        out->speakers = f->speakers.data();
        out->n_receivers = (int32_t)c.receivers.size();
        out->receivers = f->receivers.data();
        out->n_domains = (int32_t)c.domains.size();
        out->domains = f->domains.data();
        out->receiver_mode = PSTD_RECEIVER_PHYSICAL;
        return PSTD_OK;
    });
}

int pstd_file_set_scene(pstd_file *f, const pstd_config *conf) {
    return file_guarded(f, [&] {
        if (!conf) return PSTD_ERR_INVALID;
        f->data.set_scene(pstdamd::scene_conf_of(*to_conf(conf)));
        return PSTD_OK;
    });
}

int pstd_file_frame_count(pstd_file *f, int domain) {
    int n = -1;
    int rc = file_guarded(f, [&] {
        n = f->data.frame_count((uint32_t)domain);
        return PSTD_OK;
    });
    return rc ? rc : n;
}

long long pstd_file_get_frame(pstd_file *f, int frame, int domain, float *out, size_t cap) {
    long long n = -1;
    int rc = file_guarded(f, [&] {
        std::vector<float> v = f->data.frame((uint32_t)frame, (uint32_t)domain);
        n = (long long)v.size();
        if (out && cap < v.size()) {
            f->err = "output buffer smaller than the record";
            return PSTD_ERR_INVALID;
        }
        if (out) std::memcpy(out, v.data(), v.size() * sizeof(float));
        return PSTD_OK;
    });
    return rc ? rc : n;
}

long long pstd_file_get_receiver(pstd_file *f, int receiver, float *out, size_t cap) {
    long long n = -1;
    int rc = file_guarded(f, [&] {
        std::vector<float> v = f->data.receiver((uint32_t)receiver);
        n = (long long)v.size();
        if (out && cap < v.size()) {
            f->err = "output buffer smaller than the record";
            return PSTD_ERR_INVALID;
        }
        if (out) std::memcpy(out, v.data(), v.size() * sizeof(float));
        return PSTD_OK;
    });
    return rc ? rc : n;
}

long long pstd_file_get_raw(pstd_file *f, const void *key, int nkey, void *out, size_t cap) {
    long long n = -1;
    int rc = file_guarded(f, [&] {
        if (!key || nkey <= 0) return PSTD_ERR_INVALID;
        auto it = f->data.kv.records.find(std::string((const char *)key, nkey));
        if (it == f->data.kv.records.end()) {
            f->err = "no such record";
            return PSTD_ERR_INVALID;
        }
        n = (long long)it->second.size();
        if (out && cap < it->second.size()) {
            f->err = "output buffer smaller than the record";
            return PSTD_ERR_INVALID;
        }
        if (out) std::memcpy(out, it->second.data(), it->second.size());
        return PSTD_OK;
    });
    return rc ? rc : n;
}

int pstd_file_set_raw(pstd_file *f, const void *key, int nkey, const void *data, long long n, int append) {
    return file_guarded(f, [&] {
        if (!key || nkey <= 0 || n < 0 || (n > 0 && !data)) return PSTD_ERR_INVALID;
        std::string &v = f->data.kv.records[std::string((const char *)key, nkey)];
        if (!append) v.clear();
        v.append((const char *)data, (size_t)n);
        return PSTD_OK;
    });
}

int pstd_file_delete_raw(pstd_file *f, const void *key, int nkey) {
    return file_guarded(f, [&] {
        if (!key || nkey <= 0) return PSTD_ERR_INVALID;
        return f->data.kv.records.erase(std::string((const char *)key, nkey)) ? PSTD_OK : PSTD_ERR_INVALID;
    });
}

int pstd_file_record_count(pstd_file *f) {
    if (!f) return PSTD_ERR_INVALID;
    f->keys.clear();
    for (const auto &kv : f->data.kv.records) f->keys.push_back(kv.first);
    return (int)f->keys.size();
}

int pstd_file_record_key(pstd_file *f, int i, void *out, int cap) {
    if (!f || i < 0 || i >= (int)f->keys.size()) return PSTD_ERR_INVALID;
    const std::string &k = f->keys[i];
    if (out && cap < (int)k.size()) {
        f->err = "output buffer smaller than the key";
        return PSTD_ERR_INVALID;
    }
    if (out) std::memcpy(out, k.data(), k.size());
    return (int)k.size();
}

int pstd_file_commit(pstd_file *f) {
    return file_guarded(f, [&] {
        f->data.commit();
        return PSTD_OK;
    });
}

namespace {
// CLIOutput (CLI/output.cpp:30-67): frames and samples into the file, log
// lines to the caller's callback
class FileOutput : public CLog {
public:
    FileOutput(pstdamd::PSTDFileData &d, const pstd_callbacks *log) : CLog(log), data(d) {}
    void WriteFrame(int, int domain, PSTD_FRAME_PTR frame) override {
        data.save_next_frame((uint32_t)domain, frame->data(), frame->size());
    }
    void WriteSample(int, int receiver, std::vector<float> v) override {
        data.append_receiver((uint32_t)receiver, v.data(), v.size());
    }

private:
    pstdamd::PSTDFileData &data;
};
}  // namespace

int pstd_file_run(pstd_file *f, int gpu, int mcpu, int receiver_mode, const pstd_callbacks *log) {
    return file_guarded(f, [&] {
        // RunCommand::execute (CLI/main.cpp:422-452)
        pstd_config conf;
        int rc = pstd_file_get_scene(f, 0, &conf);
        if (rc) return rc;
        conf.receiver_mode = receiver_mode;
        auto out = std::make_shared<FileOutput>(f->data, log);
        out->Info("Delete old results(if any)");
        f->data.delete_results();
        out->Info("initilize new results");
        f->data.initialize_results();
        PSTDKernel kernel(gpu != 0, mcpu != 0);
        kernel.initialize_kernel(to_conf(&conf), out);
        kernel.run(out);
        f->data.commit();
        return PSTD_OK;
    });
}

}  // extern "C"
