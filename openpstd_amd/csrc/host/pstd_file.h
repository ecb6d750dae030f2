// openPSTD's scene/results file (.pstd), read and written without unqlite or
// Boost: shared/PSTDFile.{h,cpp} keeps a scene and its results as records of
// an unqlite key/value database (linear-hash engine), the scene as a Boost
// text archive of PSTDConfiguration.  This file restates
//   * the on-disk database format of unqlite 1.1.6 (the version openPSTD
//     vendors): database header page, linear-hash header page with its bucket
//     map, bucket pages with their cells and slave pages, overflow pages, the
//     DJB key hash -- the whole database is read into memory and written back
//     as one file;
//   * the text archive of PSTDConfiguration (kernel/KernelInterface.h:80-97,
//     170-205, 233-240; shared/PSTDFile.h:60-115 for QVector2D/QVector3D);
//   * PSTDFile's keys and records (shared/PSTDFile.cpp:50-58, 165-259,
//     404-492).
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "openpstd_amd.h"

namespace OpenPSTD {
namespace Kernel {
class PSTDConfiguration;
}
}  // namespace OpenPSTD

namespace pstdamd {

// unqlite's default key hash (Bernstein's h = h * 33 + c from 5381, over at
// most the first 2048 bytes).
uint32_t kv_hash(const void *key, size_t n);

// The records of an unqlite database file (key bytes -> value bytes).
class KVFile {
public:
    static constexpr uint32_t kPageSize = 4096;
    // Reads every record of the database at `path`; throws std::runtime_error
    // for a file that is not an unqlite linear-hash database.
    void load(const std::string &path);
    // Writes the records as a fresh database (page size 4096; buckets sized
    // so that each holds about one page of cells).
    void save(const std::string &path) const;
    std::map<std::string, std::string> records;
};

// The serialized content of PSTDConfiguration.
struct SceneConf {
    pstd_settings settings{};         // rk_coefficients: the first 6 of `rk`
    std::vector<float> rk;            // rk_coefficients as serialized (any length)
    std::vector<std::array<float, 3>> speakers, receivers;
    std::vector<pstd_domain_conf> domains;
};

// Boost text archive (serialization::archive) of a PSTDConfiguration.
std::string write_archive(const SceneConf &c);
SceneConf read_archive(const std::string &text);  // throws std::runtime_error

SceneConf scene_conf_of(OpenPSTD::Kernel::PSTDConfiguration &c);

// shared/PSTDFile: the records of one .pstd file.
class PSTDFileData {
public:
    static constexpr int kVersion = 4;
    enum Prefix : uint32_t {
        kScene = 1,
        kResultsScene = 100,
        kResultsSceneKernel = 101,
        kResultsFrameCount = 102,
        kResultsFrameData = 103,
        kResultsReceiverData = 104,
        kVersionKey = 10000
    };
    static std::string key(uint32_t prefix, std::initializer_list<uint32_t> args = {});

    // PSTDFile::Open: reads the file and checks its version (4).
    void open(const std::string &path);
    // PSTDFile::New: version, the default scene, an empty results scene.
    void create(const std::string &path);
    // PSTDFile::Commit: writes every record back to the file.
    void commit();

    SceneConf scene(uint32_t prefix = kScene) const;
    void set_scene(const SceneConf &c, uint32_t prefix = kScene);
    int frame_count(uint32_t domain) const;
    std::vector<float> frame(uint32_t frame, uint32_t domain) const;
    void save_next_frame(uint32_t domain, const float *data, size_t n);
    void append_receiver(uint32_t receiver, const float *data, size_t n);
    std::vector<float> receiver(uint32_t receiver) const;
    void initialize_results();
    void delete_results();

    KVFile kv;
    std::string path;

private:
    const std::string &get(const std::string &k) const;
    int get_int(const std::string &k) const;
    void set_int(const std::string &k, int v);
};

}  // namespace pstdamd
