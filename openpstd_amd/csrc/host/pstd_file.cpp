// openPSTD's .pstd scene/results file (see pstd_file.h).
#include "pstd_file.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <iterator>
#include <set>
#include <sstream>
#include <stdexcept>

#include "kernel_api.h"

namespace pstdamd {

namespace {

// ---------------------------------------------------------------- unqlite format
// Database header (page 0, unqlite.c pager_write_db_header): "unqlite", magic,
// DOS creation time, sector size, page size, KV engine name; big-endian.
constexpr char kDbSig[] = "unqlite";
constexpr uint32_t kDbMagic = 0xDB7C2712u;
// Linear-hash engine (page 1): magic, hash of a fixed word (identifies the
// hash function), free-page list, split bucket, maximum split bucket, next
// bucket-map page, map records on this page, then (logical, real) pairs.
constexpr uint32_t kHashMagic = 0xFA782DCBu;
constexpr char kHashWord[] = "chm@symisc";
constexpr uint32_t kHashHeader = 4 + 4 + 8 + 8 + 8 + 8 + 4;
constexpr uint32_t kMapPageHeader = 8 + 4;  // further map pages: next page, record count
// Bucket page: first cell offset, first free block offset, slave page.
constexpr uint32_t kPageHeader = 2 + 2 + 8;
// Cell: key hash, key length, data length, next cell offset, overflow page.
constexpr uint32_t kCellHeader = 4 + 4 + 8 + 2 + 8;
// First overflow page: next page, page and offset where the data starts.
constexpr uint32_t kOvflFirst = 8 + 8 + 2;
constexpr uint32_t kOvflNext = 8;

uint64_t be(const uint8_t *p, int n) {
    uint64_t v = 0;
    for (int i = 0; i < n; ++i) v = v << 8 | p[i];
    return v;
}
void put(uint8_t *p, uint64_t v, int n) {
    for (int i = n - 1; i >= 0; --i, v >>= 8) p[i] = (uint8_t)(v & 0xff);
}

}  // namespace

uint32_t kv_hash(const void *key, size_t n) {
    const uint8_t *p = (const uint8_t *)key;
    if (n > 2048) n = 2048;
    uint32_t h = 5381;
    for (size_t i = 0; i < n; ++i) h = h * 33 + p[i];
    return h;
}

void KVFile::load(const std::string &path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<uint8_t> buf((size_t)f.tellg());
    f.seekg(0);
    if (!f.read((char *)buf.data(), (std::streamsize)buf.size())) throw std::runtime_error("cannot read " + path);
    if (buf.size() < 512 || std::memcmp(buf.data(), kDbSig, 7) != 0 || be(&buf[7], 4) != kDbMagic)
        throw std::runtime_error(path + ": not an unqlite database");
    const uint64_t ps = be(&buf[19], 4);
    const uint64_t name_len = be(&buf[23], 2);
    if (ps < 512 || ps > 65536 || (ps & (ps - 1)) || 25 + name_len > ps ||
        std::string((const char *)&buf[25], name_len) != "hash")
        throw std::runtime_error(path + ": unsupported page size or storage engine");
    const uint64_t npages = buf.size() / ps;
    auto page = [&](uint64_t n) -> const uint8_t * {
        if (n == 0 || n >= npages) throw std::runtime_error(path + ": page number out of range");
        return &buf[n * ps];
    };
    const uint8_t *h = page(1);
    if (be(h, 4) != kHashMagic || be(h + 4, 4) != kv_hash(kHashWord, sizeof(kHashWord) - 1))
        throw std::runtime_error(path + ": not a linear-hash database");
    // bucket map: page 1 after the header, then the map chain
    std::vector<uint64_t> buckets;
    uint64_t next = be(h + 32, 8);
    uint64_t nrec = be(h + 40, 4);
    const uint8_t *rec = h + kHashHeader, *end = h + ps;
    std::set<uint64_t> seen;
    for (;;) {
        for (uint64_t r = 0; r < nrec && rec + 16 <= end; ++r, rec += 16) buckets.push_back(be(rec + 8, 8));
        if (!next) break;
        if (!seen.insert(next).second) throw std::runtime_error(path + ": bucket map loops");
        const uint8_t *m = page(next);
        next = be(m, 8);
        nrec = be(m + 8, 4);
        rec = m + kMapPageHeader;
        end = m + ps;
    }
    // One bucket: its master page and the slave pages chained from it, each a
    // chain of cells from the page header.  As unqlite's own loader
    // (lhLoadCells, lhParseOneCell): the chain is followed while each next
    // offset lies inside the page, and a cell whose payload does not check
    // out (pages that do not exist, a stored hash other than its key's) is
    // passed over -- unqlite installs such cells, and their hash never
    // matches a lookup.  unqlite leaves stale slave pointers behind (a slave
    // page later reused for other data); an offset past the page ends that
    // page's chain, and on a slave page also the bucket's slave chain, as
    // unqlite ignores a slave page that fails to load.
    struct Bad {};
    auto cell_payload = [&](const uint8_t *c, uint64_t off, std::string &key, std::string &data) {
        const uint64_t hash = be(c, 4), nkey = be(c + 4, 4), ndata = be(c + 8, 8), ovfl = be(c + 18, 8);
        if (!ovfl) {
            if (off + kCellHeader + nkey + ndata > ps) throw Bad{};
            key.assign((const char *)c + kCellHeader, nkey);
            data.assign((const char *)c + kCellHeader + nkey, ndata);
        } else {
            // key from the first overflow page on, data from the page and
            // offset that page records, both along the chain
            if (ovfl < 2 || ovfl >= npages || nkey > (1u << 30) || ndata > buf.size()) throw Bad{};
            key.reserve(nkey);
            data.reserve(ndata);
            const uint8_t *o = page(ovfl);
            const uint64_t dpage = be(o + 8, 8), doff = be(o + 16, 2);
            if (dpage < 2 || dpage >= npages || doff > ps) throw Bad{};
            uint64_t cur = ovfl, pos = kOvflFirst;
            auto step = [&]() {
                cur = be(page(cur), 8);
                if (cur < 2 || cur >= npages) throw Bad{};
                pos = kOvflNext;
            };
            for (uint64_t hops = 0; key.size() < nkey; ++hops) {
                if (hops > npages) throw Bad{};
                if (pos >= ps) step();
                const uint64_t take = std::min<uint64_t>(nkey - key.size(), ps - pos);
                key.append((const char *)page(cur) + pos, take);
                pos += take;
            }
            cur = dpage;
            pos = doff;
            for (uint64_t hops = 0; data.size() < ndata; ++hops) {
                if (hops > npages) throw Bad{};
                if (pos >= ps) step();
                const uint64_t take = std::min<uint64_t>(ndata - data.size(), ps - pos);
                data.append((const char *)page(cur) + pos, take);
                pos += take;
            }
        }
        if (kv_hash(key.data(), key.size()) != hash) throw Bad{};
    };
    struct Cell {
        uint32_t hash;
        bool valid;
        std::string key, data;
    };
    // returns false when the page's chain leaves the page
    auto parse_page = [&](uint64_t pn, std::vector<Cell> &out, uint64_t &slave) {
        const uint8_t *p = page(pn);
        slave = be(p + 4, 8);
        uint64_t off = be(p, 2);
        for (uint64_t guard = 0; off; ++guard) {
            if (off + kCellHeader > ps || guard > ps) return false;
            const uint8_t *c = p + off;
            const uint64_t cnext = be(c + 16, 2);
            if (cnext >= ps) return false;
            Cell cell{(uint32_t)be(c, 4), true, {}, {}};
            try {
                cell_payload(c, off, cell.key, cell.data);
            } catch (const Bad &) {
                cell.valid = false;
                cell.key.clear();
                cell.data.clear();
            }
            out.push_back(std::move(cell));
            off = cnext;
        }
        return true;
    };
    // Which cell a lookup finds when a key occurs more than once in a bucket
    // (stale copies): unqlite installs a bucket's cells, in load order, at the
    // front of the collision chains of a table of 32 chains that doubles
    // whenever it holds 3 cells per chain (below 100000 cells), re-inserting
    // from the newest cell to the oldest (lhInstallCell); a lookup takes the
    // first matching cell of the key's chain (lhFindCell).
    auto resolve = [&](std::vector<Cell> &cells) {
        size_t size = 32;
        std::vector<std::deque<size_t>> table(size);
        for (size_t i = 0; i < cells.size(); ++i) {
            table[cells[i].hash & (size - 1)].push_front(i);
            const size_t n = i + 1;
            if (n >= size * 3 && n < 100000) {
                size <<= 1;
                std::vector<std::deque<size_t>> grown(size);
                for (size_t j = n; j-- > 0;) grown[cells[j].hash & (size - 1)].push_front(j);
                table.swap(grown);
            }
        }
        for (const auto &chain : table)
            for (size_t i : chain)
                if (cells[i].valid) records.emplace(cells[i].key, std::move(cells[i].data));
    };
    records.clear();
    std::set<uint64_t> masters(buckets.begin(), buckets.end());
    for (uint64_t master : buckets) {
        std::vector<Cell> cells;
        uint64_t slave = 0;
        if (!parse_page(master, cells, slave))
            throw std::runtime_error(path + ": corrupt bucket page " + std::to_string(master));
        std::set<uint64_t> chain{master};
        while (slave) {
            if (slave >= npages || masters.count(slave) || !chain.insert(slave).second) break;
            if (!parse_page(slave, cells, slave)) break;
        }
        resolve(cells);
    }
}

void KVFile::save(const std::string &path) const {
    const uint32_t ps = kPageSize;
    struct Rec {
        const std::string *k, *v;
        uint32_t hash;
        bool local;
    };
    std::vector<Rec> recs;
    uint64_t bytes = 0;
    for (const auto &kv : records) {
        const uint64_t sz = kCellHeader + kv.first.size() + kv.second.size();
        const bool local = sz <= ps / 2;
        recs.push_back({&kv.first, &kv.second, kv_hash(kv.first.data(), kv.first.size()), local});
        bytes += local ? sz : kCellHeader;
    }
    // buckets (a power of two) for about 3/4 of a page of cells each; split
    // bucket 0, so a key's bucket is hash & (buckets - 1)
    uint64_t nb = 1;
    while (nb * (ps - kPageHeader) * 3 / 4 < bytes) nb <<= 1;
    std::vector<std::vector<uint8_t>> pages(2, std::vector<uint8_t>(ps, 0));
    auto new_page = [&]() {
        pages.emplace_back(ps, 0);
        return (uint64_t)pages.size() - 1;
    };
    // bucket map pages after page 1
    const uint64_t on_first = (ps - kHashHeader) / 16, per_map = (ps - kMapPageHeader) / 16;
    std::vector<uint64_t> map_pages;
    for (uint64_t left = nb > on_first ? nb - on_first : 0; left > 0; left -= std::min(left, per_map))
        map_pages.push_back(new_page());
    std::vector<std::vector<const Rec *>> bucket(nb);
    for (const Rec &r : recs) bucket[r.hash & (nb - 1)].push_back(&r);
    auto write_overflow = [&](const std::string &k, const std::string &v) {
        const uint64_t first = new_page();
        uint64_t cur = first, pos = kOvflFirst;
        auto write = [&](const std::string &s) {
            size_t done = 0;
            while (done < s.size()) {
                if (pos >= ps) {
                    const uint64_t np = new_page();
                    put(pages[cur].data(), np, 8);
                    cur = np;
                    pos = kOvflNext;
                }
                const size_t take = std::min<size_t>(s.size() - done, ps - pos);
                std::memcpy(pages[cur].data() + pos, s.data() + done, take);
                done += take;
                pos += take;
            }
        };
        write(k);
        put(pages[first].data() + 8, cur, 8);
        put(pages[first].data() + 16, pos, 2);
        write(v);
        return first;
    };
    std::vector<uint64_t> real(nb);
    for (uint64_t b = 0; b < nb; ++b) {
        uint64_t pg = new_page();
        real[b] = pg;
        uint32_t used = kPageHeader;
        auto close_page = [&]() {
            if (ps - used > 3) {  // the rest of the page as one free block
                put(pages[pg].data() + used, 0, 2);
                put(pages[pg].data() + used + 2, ps - used, 2);
                put(pages[pg].data() + 2, used, 2);
            }
        };
        for (const Rec *r : bucket[b]) {
            const uint32_t sz = r->local ? kCellHeader + (uint32_t)(r->k->size() + r->v->size()) : kCellHeader;
            if (used + sz > ps) {
                close_page();
                const uint64_t slave = new_page();
                put(pages[pg].data() + 4, slave, 8);
                pg = slave;
                used = kPageHeader;
            }
            const uint64_t ovfl = r->local ? 0 : write_overflow(*r->k, *r->v);
            uint8_t *c = pages[pg].data() + used;
            put(c, r->hash, 4);
            put(c + 4, r->k->size(), 4);
            put(c + 8, r->v->size(), 8);
            put(c + 16, be(pages[pg].data(), 2), 2);  // previous first cell
            put(c + 18, ovfl, 8);
            if (r->local) {
                std::memcpy(c + kCellHeader, r->k->data(), r->k->size());
                std::memcpy(c + kCellHeader + r->k->size(), r->v->data(), r->v->size());
            }
            put(pages[pg].data(), used, 2);  // first cell
            used += sz;
        }
        close_page();
    }
    // page 0: database header
    uint8_t *p0 = pages[0].data();
    std::memcpy(p0, kDbSig, 7);
    put(p0 + 7, kDbMagic, 4);
    put(p0 + 11, 0, 4);    // creation time (DOS format), not recorded
    put(p0 + 15, 512, 4);  // sector size
    put(p0 + 19, ps, 4);
    put(p0 + 23, 4, 2);
    std::memcpy(p0 + 25, "hash", 4);
    // page 1: linear-hash header and the bucket map
    uint8_t *h = pages[1].data();
    put(h, kHashMagic, 4);
    put(h + 4, kv_hash(kHashWord, sizeof(kHashWord) - 1), 4);
    put(h + 8, 0, 8);    // free pages
    put(h + 16, 0, 8);   // split bucket
    put(h + 24, nb, 8);  // maximum split bucket
    uint64_t b = 0;
    auto fill_map = [&](uint8_t *pg, uint32_t at, uint64_t count) {
        for (uint64_t i = 0; i < count; ++i, ++b, at += 16) {
            put(pg + at, b, 8);
            put(pg + at + 8, real[b], 8);
        }
    };
    const uint64_t first_n = std::min(nb, on_first);
    put(h + 32, map_pages.empty() ? 0 : map_pages[0], 8);
    put(h + 40, first_n, 4);
    fill_map(h, kHashHeader, first_n);
    for (size_t m = 0; m < map_pages.size(); ++m) {
        uint8_t *mp = pages[map_pages[m]].data();
        const uint64_t n = std::min(per_map, nb - b);
        put(mp, m + 1 < map_pages.size() ? map_pages[m + 1] : 0, 8);
        put(mp + 8, n, 4);
        fill_map(mp, kMapPageHeader, n);
    }
    const std::string tmp = path + ".tmp";
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        if (!f) throw std::runtime_error("cannot write " + tmp);
        for (const auto &pg : pages) f.write((const char *)pg.data(), ps);
        if (!f) throw std::runtime_error("write failed: " + tmp);
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) throw std::runtime_error("cannot replace " + path);
}

// ---------------------------------------------------------------- text archive
// boost::archive::text_oarchive of PSTDConfiguration: a header
// "22 serialization::archive <library version>", then the values in the
// order the serialize() functions list them, space separated.  Every class
// type (PSTDConfiguration, PSTDSettings, std::vector<QVector3D>, QVector3D,
// std::vector<DomainConf>, DomainConf, QVector2D, DomainConfEdge) is preceded
// at its first occurrence by its tracking level and class version (0 0);
// vectors write their size and the item version (0) before the items; a
// vector of a primitive type (rk_coefficients) has no class preamble; bools
// are 0/1.
namespace {

constexpr int kArchiveLibraryVersion = 12;  // written; any version >= 4 is read

struct Writer {
    std::ostringstream o;
    void tok(const std::string &s) { o << ' ' << s; }
    void f(float v) {
        char b[32];
        std::snprintf(b, sizeof b, "%.9g", (double)v);
        tok(b);
    }
    void i(long long v) { tok(std::to_string(v)); }
    void pre() { tok("0"), tok("0"); }
};

struct Reader {
    std::vector<std::string> t;
    size_t at = 0;
    explicit Reader(const std::string &s) {
        std::istringstream in(s);
        std::string w;
        while (in >> w) t.push_back(w);
    }
    const std::string &next() {
        if (at >= t.size()) throw std::runtime_error("scene archive ends early");
        return t[at++];
    }
    float f() {
        const std::string &s = next();
        char *e = nullptr;
        const float v = std::strtof(s.c_str(), &e);
        if (!e || *e) throw std::runtime_error("scene archive: not a number: " + s);
        return v;
    }
    long long i() {
        const std::string &s = next();
        char *e = nullptr;
        const long long v = std::strtoll(s.c_str(), &e, 10);
        if (!e || *e) throw std::runtime_error("scene archive: not an integer: " + s);
        return v;
    }
    void pre() {
        const long long tracking = i(), version = i();
        if (tracking != 0 && tracking != 1) throw std::runtime_error("scene archive: bad tracking level");
        if (version < 0) throw std::runtime_error("scene archive: bad class version");
    }
    size_t count() {
        const long long n = i();
        if (n < 0 || n > (1 << 24)) throw std::runtime_error("scene archive: bad collection size");
        return (size_t)n;
    }
};

void parse(Reader &r, SceneConf &c, bool vector_preamble) {
    if (r.next() != "22" || r.next() != "serialization::archive")
        throw std::runtime_error("not a Boost text archive");
    const long long lib = r.i();
    if (lib < 4) throw std::runtime_error("Boost archive library version below 4 is not supported");
    r.pre();  // PSTDConfiguration
    r.pre();  // PSTDSettings
    pstd_settings &s = c.settings;
    float *fl[] = {&s.calctime, &s.c1, &s.ampmax, &s.freqmax, &s.rho, &s.patcherror, &s.tfactRK, &s.grid_spacing,
                   &s.band_width, &s.wave_length};
    for (float *p : fl) *p = r.f();
    c.rk.resize(r.count());
    r.i();  // item version
    for (float &v : c.rk) v = r.f();
    for (int k = 0; k < 6; ++k) s.rk_coefficients[k] = k < (int)c.rk.size() ? c.rk[k] : 0.f;
    s.spectral_interpolation = (int32_t)r.i();
    s.pml_cells = (int32_t)r.i();
    s.save_nth = (int32_t)r.i();
    bool v3_seen = false;
    for (auto *vec : {&c.speakers, &c.receivers}) {
        if (vector_preamble && vec == &c.speakers) r.pre();
        vec->resize(r.count());
        r.i();
        for (auto &p : *vec) {
            if (!v3_seen) r.pre(), v3_seen = true;
            p = {r.f(), r.f(), r.f()};
        }
    }
    if (vector_preamble) r.pre();
    c.domains.resize(r.count());
    r.i();
    bool first = true;
    for (pstd_domain_conf &d : c.domains) {
        if (first) r.pre();  // DomainConf
        if (first) r.pre();  // QVector2D
        d.top_left[0] = r.f();
        d.top_left[1] = r.f();
        d.size[0] = r.f();
        d.size[1] = r.f();
        for (pstd_edge *e : {&d.T, &d.L, &d.B, &d.R}) {
            if (first && e == &d.T) r.pre();  // DomainConfEdge
            e->absorption = r.f();
            e->lr = (int32_t)r.i();
        }
        first = false;
    }
    if (r.at != r.t.size()) throw std::runtime_error("scene archive: trailing values");
}

}  // namespace

std::string write_archive(const SceneConf &c) {
    Writer w;
    w.o << "22 serialization::archive " << kArchiveLibraryVersion;
    w.pre();  // PSTDConfiguration
    w.pre();  // PSTDSettings
    const pstd_settings &s = c.settings;
    for (float v : {s.calctime, s.c1, s.ampmax, s.freqmax, s.rho, s.patcherror, s.tfactRK, s.grid_spacing,
                    s.band_width, s.wave_length})
        w.f(v);
    w.i((long long)c.rk.size());
    w.i(0);
    for (float v : c.rk) w.f(v);
    w.i(s.spectral_interpolation ? 1 : 0);
    w.i(s.pml_cells);
    w.i(s.save_nth);
    bool v3_seen = false;
    w.pre();  // std::vector<QVector3D>
    for (const auto *vec : {&c.speakers, &c.receivers}) {
        w.i((long long)vec->size());
        w.i(0);
        for (const auto &p : *vec) {
            if (!v3_seen) w.pre(), v3_seen = true;
            w.f(p[0]), w.f(p[1]), w.f(p[2]);
        }
    }
    w.pre();  // std::vector<DomainConf>
    w.i((long long)c.domains.size());
    w.i(0);
    bool first = true;
    for (const pstd_domain_conf &d : c.domains) {
        if (first) w.pre(), w.pre();  // DomainConf, QVector2D
        w.f(d.top_left[0]), w.f(d.top_left[1]), w.f(d.size[0]), w.f(d.size[1]);
        for (const pstd_edge *e : {&d.T, &d.L, &d.B, &d.R}) {
            if (first && e == &d.T) w.pre();  // DomainConfEdge
            w.f(e->absorption);
            w.i(e->lr ? 1 : 0);
        }
        first = false;
    }
    w.o << '\n';
    return w.o.str();
}

SceneConf read_archive(const std::string &text) {
    // the class preambles of the two vectors of class type, as written above;
    // an archive without them is accepted too
    std::string first_error;
    for (bool vp : {true, false}) {
        try {
            Reader r(text);
            SceneConf c;
            parse(r, c, vp);
            return c;
        } catch (const std::exception &e) {
            if (first_error.empty()) first_error = e.what();
        }
    }
    throw std::runtime_error(first_error);
}

SceneConf scene_conf_of(OpenPSTD::Kernel::PSTDConfiguration &conf) {
    using namespace OpenPSTD::Kernel;
    SceneConf c;
    PSTDSettings &st = conf.Settings;
    pstd_settings &s = c.settings;
    s.calctime = st.GetRenderTime();
    s.c1 = st.GetSoundSpeed();
    s.ampmax = st.GetAttenuationOfPMLCells();
    s.freqmax = st.GetMaxFrequency();
    s.rho = st.GetDensityOfAir();
    s.patcherror = st.GetPatchError();
    s.tfactRK = st.GetFactRK();
    s.grid_spacing = st.GetGridSpacing();
    s.band_width = st.GetBandWidth();
    s.wave_length = st.GetWaveLength();
    c.rk = st.GetRKCoefficients();
    for (int k = 0; k < 6; ++k) s.rk_coefficients[k] = k < (int)c.rk.size() ? c.rk[k] : 0.f;
    s.spectral_interpolation = st.GetSpectralInterpolation();
    s.pml_cells = st.GetPMLCells();
    s.save_nth = st.GetSaveNth();
    for (const QVector3D &p : conf.Speakers) c.speakers.push_back({p.x(), p.y(), p.z()});
    for (const QVector3D &p : conf.Receivers) c.receivers.push_back({p.x(), p.y(), p.z()});
    for (DomainConf &d : conf.Domains) {
        pstd_domain_conf o{};
        o.top_left[0] = d.TopLeft.x();
        o.top_left[1] = d.TopLeft.y();
        o.size[0] = d.Size.x();
        o.size[1] = d.Size.y();
        o.T = {d.T.Absorption, d.T.LR};
        o.L = {d.L.Absorption, d.L.LR};
        o.B = {d.B.Absorption, d.B.LR};
        o.R = {d.R.Absorption, d.R.LR};
        c.domains.push_back(o);
    }
    return c;
}

// ---------------------------------------------------------------- PSTDFile
std::string PSTDFileData::key(uint32_t prefix, std::initializer_list<uint32_t> args) {
    // PSTDFile::CreateKey (shared/PSTDFile.cpp:344-366): native 32-bit ints
    std::string k((const char *)&prefix, 4);
    for (uint32_t a : args) k.append((const char *)&a, 4);
    return k;
}

const std::string &PSTDFileData::get(const std::string &k) const {
    auto it = kv.records.find(k);
    if (it == kv.records.end()) throw std::runtime_error("no record " + std::to_string(*(const uint32_t *)k.data()));
    return it->second;
}

int PSTDFileData::get_int(const std::string &k) const {
    const std::string &v = get(k);
    if (v.size() < sizeof(int)) throw std::runtime_error("integer record too short");
    int x;
    std::memcpy(&x, v.data(), sizeof x);
    return x;
}

void PSTDFileData::set_int(const std::string &k, int v) { kv.records[k] = std::string((const char *)&v, sizeof v); }

void PSTDFileData::open(const std::string &p) {
    kv.load(p);
    path = p;
    const int version = get_int(key(kVersionKey));
    if (version != kVersion)
        throw std::runtime_error("Wrong version: " + std::to_string(version) + "(expected: " +
                                 std::to_string(kVersion) + ")");
}

void PSTDFileData::create(const std::string &p) {
    kv.records.clear();
    path = p;
    set_int(key(kVersionKey), kVersion);
    set_scene(scene_conf_of(*OpenPSTD::Kernel::PSTDConfiguration::CreateDefaultConf()));
    set_scene(scene_conf_of(*OpenPSTD::Kernel::PSTDConfiguration::CreateEmptyConf()), kResultsScene);
    commit();
}

void PSTDFileData::commit() { kv.save(path); }

SceneConf PSTDFileData::scene(uint32_t prefix) const { return read_archive(get(key(prefix))); }

void PSTDFileData::set_scene(const SceneConf &c, uint32_t prefix) { kv.records[key(prefix)] = write_archive(c); }

int PSTDFileData::frame_count(uint32_t domain) const { return get_int(key(kResultsFrameCount, {domain})); }

std::vector<float> PSTDFileData::frame(uint32_t fr, uint32_t domain) const {
    const std::string &v = get(key(kResultsFrameData, {domain, fr}));
    std::vector<float> out(v.size() / 4);
    std::memcpy(out.data(), v.data(), out.size() * 4);
    return out;
}

void PSTDFileData::save_next_frame(uint32_t domain, const float *data, size_t n) {
    // PSTDFile::SaveNextResultsFrame / IncrementFrameCount
    const int fr = frame_count(domain);
    set_int(key(kResultsFrameCount, {domain}), fr + 1);
    kv.records[key(kResultsFrameData, {domain, (uint32_t)fr})] = std::string((const char *)data, n * 4);
}

void PSTDFileData::append_receiver(uint32_t r, const float *data, size_t n) {
    kv.records[key(kResultsReceiverData, {r})].append((const char *)data, n * 4);
}

std::vector<float> PSTDFileData::receiver(uint32_t r) const {
    const std::string &v = get(key(kResultsReceiverData, {r}));
    std::vector<float> out(v.size() / 4);
    std::memcpy(out.data(), v.data(), out.size() * 4);
    return out;
}

void PSTDFileData::initialize_results() {
    // PSTDFile::InitializeResults (shared/PSTDFile.cpp:228-243)
    const SceneConf c = scene();
    set_scene(c, kResultsScene);
    for (uint32_t i = 0; i < c.domains.size(); ++i) set_int(key(kResultsFrameCount, {i}), 0);
    for (uint32_t i = 0; i < c.receivers.size(); ++i) kv.records[key(kResultsReceiverData, {i})] = std::string();
}

void PSTDFileData::delete_results() {
    // PSTDFile::DeleteResults (shared/PSTDFile.cpp:245-259)
    const SceneConf rc = scene(kResultsScene);
    for (uint32_t d = 0; d < rc.domains.size(); ++d) {
        auto it = kv.records.find(key(kResultsFrameCount, {d}));
        if (it == kv.records.end()) continue;
        const int n = get_int(it->first);
        for (int f = 0; f < n; ++f) kv.records.erase(key(kResultsFrameData, {d, (uint32_t)f}));
        kv.records.erase(key(kResultsFrameCount, {d}));
    }
    set_scene(scene_conf_of(*OpenPSTD::Kernel::PSTDConfiguration::CreateEmptyConf()), kResultsScene);
}

}  // namespace pstdamd
