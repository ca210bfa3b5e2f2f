// TIME snapshots of a MeasurementSet main table from its columns exported as NumPy .npy files
// (casacore is not part of this build): the reader of the integrated CLI,
// read_measurement_set_snapshot / group_rows_by_time / read_measurement_set_rows
// (SVDIntegratedImager.cu:180-378), with the casacore table layer replaced by one file per
// column.
//
// A column directory holds, as python-casacore's getcol returns them (C order, the row axis
// first, each row's axes in the reverse of casacore's IPosition order):
//   TIME.npy                       float64 [rows]
//   UVW.npy                        float64 [rows, 3]
//   DATA.npy                       complex64 [rows, a, b]   (IPosition shape of a row: (b, a))
//   FLAG.npy                       bool [rows, a, b]
//   WEIGHT_SPECTRUM.npy            float32 [rows, a, b]     (optional: weights 1)
//   SPECTRAL_WINDOW/CHAN_FREQ.npy  float64 [spectral windows, channels] or [channels]
//   OBSERVATION/ARRAY_CENTER.npy   float64 [observations, k] or [k]   (optional: 0)
// Rows are grouped by exact TIME value in ascending order (a std::map, as the reference does);
// a snapshot's rows keep their table order.  Per row, the channel and polarisation axes follow
// detect_row_axis_layout (SVDIntegratedImager.cu:99-134) on the IPosition shape, polarisations
// 0 and 3 are taken, and channel c of the concatenated CHAN_FREQ list reads the row's channel
// c mod (its channel count) (:299-322).  UVW and CHAN_FREQ stay FP64 (the reference narrows
// them to float).
#include "msreader.hpp"

#include <cerrno>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iterator>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include <sys/types.h>

namespace svdi {
namespace {

enum class Kind { F64, F32, C64, C128, B1, U1 };

// A .npy file's element kind, shape (file_rows: the length of its first axis in the file) and
// raw little-endian data: the whole array, or only some of its rows (the first axis), in the
// order asked, shape[0] then being their count.
struct NpyArray {
    Kind kind = Kind::F64;
    std::vector<size_t> shape;
    size_t file_rows = 0;
    std::vector<char> data;
    size_t count() const {
        size_t n = 1;
        for (size_t d : shape) n *= d;
        return n;
    }
};

size_t kind_size(Kind k) {
    switch (k) {
        case Kind::F64: return 8;
        case Kind::F32: return 4;
        case Kind::C64: return 8;
        case Kind::C128: return 16;
        default: return 1;
    }
}

// The value of `key` in the header dict (the text after "'key':"), or throws.
std::string header_field(const std::string& header, const std::string& key, const std::string& path) {
    const std::string tag = "'" + key + "'";
    const size_t k = header.find(tag);
    if (k == std::string::npos) throw std::runtime_error(path + ": .npy header has no " + key);
    size_t p = header.find(':', k + tag.size());
    if (p == std::string::npos) throw std::runtime_error(path + ": malformed .npy header");
    ++p;
    while (p < header.size() && header[p] == ' ') ++p;
    return header.substr(p);
}

struct FileCloser {
    void operator()(std::FILE* f) const { std::fclose(f); }
};

void read_exact(std::FILE* f, void* dst, size_t bytes, off_t at, const std::string& path) {
    if (bytes == 0) return;
    if (fseeko(f, at, SEEK_SET) != 0 || std::fread(dst, 1, bytes, f) != bytes)
        throw std::runtime_error("cannot read " + path + " (truncated?)");
}

// `rows`: the rows (first-axis indices) to read, or null for the whole array.  A table column
// can be far larger than one snapshot's rows, which are read run by run of consecutive rows.
NpyArray read_npy(const std::string& path, const std::vector<size_t>* rows = nullptr) {
    std::unique_ptr<std::FILE, FileCloser> f(std::fopen(path.c_str(), "rb"));
    if (!f) throw std::runtime_error("cannot open " + path + ": " + std::strerror(errno));
    unsigned char pre[12] = {};
    const size_t got = std::fread(pre, 1, sizeof pre, f.get());
    if (got < 10 || std::memcmp(pre, "\x93NUMPY", 6) != 0) throw std::runtime_error(path + " is not a .npy file");
    const unsigned major = pre[6];
    size_t hlen, hoff;
    if (major == 1) {
        hlen = pre[8] | (static_cast<size_t>(pre[9]) << 8);
        hoff = 10;
    } else if (major == 2 || major == 3) {
        if (got < 12) throw std::runtime_error(path + " is truncated");
        hlen = 0;
        for (int i = 0; i < 4; ++i) hlen |= static_cast<size_t>(pre[8 + i]) << (8 * i);
        hoff = 12;
    } else {
        throw std::runtime_error(path + ": unsupported .npy version " + std::to_string(major));
    }
    if (hlen > (1u << 20)) throw std::runtime_error(path + ": .npy header too long");
    std::string header(hlen, ' ');
    read_exact(f.get(), &header[0], hlen, static_cast<off_t>(hoff), path);

    NpyArray a;
    const std::string descr = header_field(header, "descr", path);
    struct { const char* text; Kind kind; } kinds[] = {
        {"'<f8'", Kind::F64}, {"'<f4'", Kind::F32}, {"'<c8'", Kind::C64}, {"'<c16'", Kind::C128},
        {"'|b1'", Kind::B1}, {"'|u1'", Kind::U1}};
    bool known = false;
    for (const auto& k : kinds) {
        if (descr.compare(0, std::strlen(k.text), k.text) == 0) {
            a.kind = k.kind;
            known = true;
            break;
        }
    }
    if (!known) throw std::runtime_error(path + ": unsupported element type " + descr.substr(0, descr.find(',')));
    if (header_field(header, "fortran_order", path).compare(0, 5, "False") != 0)
        throw std::runtime_error(path + ": Fortran-ordered arrays are not supported (save C order)");
    const std::string shape = header_field(header, "shape", path);
    if (shape.empty() || shape[0] != '(') throw std::runtime_error(path + ": malformed shape");
    const size_t close = shape.find(')');
    if (close == std::string::npos) throw std::runtime_error(path + ": malformed shape");
    size_t p = 1;
    while (p < close) {
        while (p < close && (shape[p] == ' ' || shape[p] == ',')) ++p;
        if (p >= close) break;
        size_t used = 0;
        const unsigned long long v = std::stoull(shape.substr(p, close - p), &used);
        a.shape.push_back(static_cast<size_t>(v));
        p += used;
    }
    const size_t esize = kind_size(a.kind);
    const off_t doff = static_cast<off_t>(hoff + hlen);
    if (fseeko(f.get(), 0, SEEK_END) != 0) throw std::runtime_error("cannot read " + path);
    const off_t fsize = ftello(f.get());
    if (fsize < doff || static_cast<size_t>(fsize - doff) < a.count() * esize)
        throw std::runtime_error(path + " holds fewer elements than its shape");
    if (rows == nullptr) {
        a.file_rows = a.shape.empty() ? 1 : a.shape[0];
        a.data.resize(a.count() * esize);
        read_exact(f.get(), a.data.data(), a.data.size(), doff, path);
        return a;
    }
    if (a.shape.empty()) throw std::runtime_error(path + ": expected one entry per row");
    a.file_rows = a.shape[0];
    const size_t row_bytes = (a.file_rows ? a.count() / a.file_rows : 0) * esize;
    a.shape[0] = rows->size();
    a.data.resize(rows->size() * row_bytes);
    for (size_t i = 0; i < rows->size();) {
        size_t j = i + 1;
        while (j < rows->size() && (*rows)[j] == (*rows)[j - 1] + 1) ++j;
        if ((*rows)[j - 1] >= a.file_rows) throw std::runtime_error(path + " has fewer rows than TIME");
        read_exact(f.get(), a.data.data() + i * row_bytes, (j - i) * row_bytes,
                   doff + static_cast<off_t>((*rows)[i] * row_bytes), path);
        i = j;
    }
    return a;
}

bool file_exists(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fclose(f);
    return true;
}

double real_at(const NpyArray& a, size_t i) {
    switch (a.kind) {
        case Kind::F64: { double v; std::memcpy(&v, a.data.data() + 8 * i, 8); return v; }
        case Kind::F32: { float v; std::memcpy(&v, a.data.data() + 4 * i, 4); return v; }
        default: throw std::logic_error("not a real array");
    }
}

void require_real(const NpyArray& a, const std::string& name) {
    if (a.kind != Kind::F64 && a.kind != Kind::F32) throw std::runtime_error(name + " must hold float64 or float32");
}

// Element i of a complex column as complex<float> (the reference's casacore::Complex).
std::complex<float> complex_at(const NpyArray& a, size_t i) {
    if (a.kind == Kind::C64) {
        float v[2];
        std::memcpy(v, a.data.data() + 8 * i, 8);
        return {v[0], v[1]};
    }
    double v[2];
    std::memcpy(v, a.data.data() + 16 * i, 16);
    return {static_cast<float>(v[0]), static_cast<float>(v[1])};
}

// detect_row_axis_layout (SVDIntegratedImager.cu:99-134) on a row's IPosition shape (d0, d1).
struct RowLayout {
    int channel_axis, pol_axis;
    size_t num_channels, num_pols;
};
RowLayout row_layout(size_t d0, size_t d1) {
    auto small = [](size_t d) { return d == 4 || d == 2 || d == 1; };
    if (small(d1)) return {0, 1, d0, d1};
    if (small(d0)) return {1, 0, d1, d0};
    if (d0 >= d1) return {0, 1, d0, d1};
    return {1, 0, d1, d0};
}

}  // namespace

MsSnapshot read_ms_snapshot(const std::string& dir, size_t snapshot_index) {
    const std::string base = dir.empty() || dir.back() == '/' ? dir : dir + "/";
    // group_rows_by_time (SVDIntegratedImager.cu:333-357)
    const NpyArray time = read_npy(base + "TIME.npy");
    require_real(time, "TIME");
    if (time.shape.size() != 1) throw std::runtime_error("TIME must be one value per row");
    const size_t table_rows = time.shape[0];
    std::map<double, std::vector<size_t>> groups;
    for (size_t r = 0; r < table_rows; ++r) groups[real_at(time, r)].push_back(r);

    MsSnapshot s;
    s.total_snapshots = groups.size();
    // read_measurement_set_snapshot (SVDIntegratedImager.cu:359-378)
    if (snapshot_index >= groups.size())
        throw std::runtime_error("Snapshot_Index is outside available TIME snapshots");
    auto g = groups.begin();
    std::advance(g, static_cast<std::ptrdiff_t>(snapshot_index));
    s.index = snapshot_index;
    s.time = g->first;
    const std::vector<size_t>& rows = g->second;

    // read_measurement_set_rows (SVDIntegratedImager.cu:180-325)
    // Only the snapshot's rows of the per-row columns are read.
    const NpyArray uvw = read_npy(base + "UVW.npy", &rows);
    const NpyArray data = read_npy(base + "DATA.npy", &rows);
    const NpyArray flag = read_npy(base + "FLAG.npy", &rows);
    const bool has_weight = file_exists(base + "WEIGHT_SPECTRUM.npy");
    NpyArray weight;
    if (has_weight) weight = read_npy(base + "WEIGHT_SPECTRUM.npy", &rows);
    require_real(uvw, "UVW");
    if (uvw.shape.size() != 2 || uvw.file_rows != table_rows || uvw.shape[1] != 3)
        throw std::runtime_error("UVW does not have 3 entries per row of TIME");
    if (data.kind != Kind::C64 && data.kind != Kind::C128) throw std::runtime_error("DATA must hold complex values");
    if (data.shape.size() != 3) throw std::runtime_error("DATA does not have 2 dimensions per row");
    if (data.file_rows != table_rows) throw std::runtime_error("DATA does not have one row per row of TIME");
    if (flag.kind != Kind::B1 && flag.kind != Kind::U1) throw std::runtime_error("FLAG must hold booleans");
    if (flag.shape != data.shape || flag.file_rows != table_rows) throw std::runtime_error("FLAG's shape differs from DATA's");
    if (has_weight) {
        require_real(weight, "WEIGHT_SPECTRUM");
        if (weight.shape != data.shape || weight.file_rows != table_rows)
            throw std::runtime_error("WEIGHT_SPECTRUM's shape differs from DATA's");
    }

    const NpyArray chan_freq = read_npy(base + "SPECTRAL_WINDOW/CHAN_FREQ.npy");
    require_real(chan_freq, "CHAN_FREQ");
    if (chan_freq.shape.empty() || chan_freq.shape.size() > 2)
        throw std::runtime_error("CHAN_FREQ must be [spectral windows, channels] or [channels]");
    for (size_t i = 0; i < chan_freq.count(); ++i) s.freq_hz.push_back(real_at(chan_freq, i));
    if (s.freq_hz.empty()) throw std::runtime_error("No frequencies found in SPECTRAL_WINDOW for " + dir);

    // ARRAY_CENTER x of the first OBSERVATION row, 0 when absent or unreadable (:205-218).
    s.array_center_x = 0.0;
    const std::string ac_path = base + "OBSERVATION/ARRAY_CENTER.npy";
    if (file_exists(ac_path)) {
        try {
            const NpyArray ac = read_npy(ac_path);
            if ((ac.kind == Kind::F64 || ac.kind == Kind::F32) && ac.count() > 0 && !ac.shape.empty() &&
                ac.shape[0] > 0)
                s.array_center_x = real_at(ac, 0);
        } catch (const std::exception&) {
            s.array_center_x = 0.0;
        }
    }

    const size_t R = rows.size();
    const size_t C = s.freq_hz.size();
    s.num_rows = R;
    s.num_channels = C;
    const size_t S = R * C;
    s.uvw.resize(3 * R);
    s.vis0_re.resize(S);
    s.vis0_im.resize(S);
    s.vis3_re.resize(S);
    s.vis3_im.resize(S);
    s.flag0.resize(S);
    s.flag3.resize(S);
    s.has_weights = has_weight;
    if (has_weight) {
        s.weight0.resize(S);
        s.weight3.resize(S);
    }
    // A row of DATA is a = shape[1] by b = shape[2] in C order: IPosition shape (b, a), element
    // (i0, i1) at i1 * b + i0.
    const size_t a = data.shape[1], b = data.shape[2];
    const RowLayout L = row_layout(b, a);
    if (L.num_pols < 4) throw std::runtime_error("DATA has fewer than 4 polarisations in " + dir);
    if (L.num_channels == 0) throw std::runtime_error("DATA has zero channels in " + dir);
    auto element = [&](size_t chan, size_t pol) {
        size_t ip[2];
        ip[L.channel_axis] = chan;
        ip[L.pol_axis] = pol;
        return ip[1] * b + ip[0];
    };
    const size_t per_row = a * b;
    for (size_t lr = 0; lr < R; ++lr) {                    // (the arrays hold the snapshot's rows)
        for (int k = 0; k < 3; ++k) s.uvw[3 * lr + k] = real_at(uvw, 3 * lr + k);
        for (size_t c = 0; c < C; ++c) {
            const size_t dst = lr + c * R;
            const size_t cr = c % L.num_channels;
            const size_t e0 = lr * per_row + element(cr, 0), e3 = lr * per_row + element(cr, 3);
            const std::complex<float> v0 = complex_at(data, e0), v3 = complex_at(data, e3);
            s.vis0_re[dst] = v0.real();
            s.vis0_im[dst] = v0.imag();
            s.vis3_re[dst] = v3.real();
            s.vis3_im[dst] = v3.imag();
            s.flag0[dst] = flag.data[e0] != 0 ? 1 : 0;
            s.flag3[dst] = flag.data[e3] != 0 ? 1 : 0;
            if (has_weight) {
                s.weight0[dst] = static_cast<float>(real_at(weight, e0));
                s.weight3[dst] = static_cast<float>(real_at(weight, e3));
            }
        }
    }
    return s;
}

}  // namespace svdi
