// lt_io.cpp -- files exchanged with the reference: BMW gauge configurations
// (LettuceGauge/IO/config_io/bmw_format.hpp) and the metadynamics bias potential
// (LettuceGauge/actions/bias_potential/metadynamics.hpp:304-505).
//
// Configurations are packed and unpacked on the device (lt_bmw.hip) in slabs of
// whole timeslices' worth of BMW sites; the host only moves bytes between the
// file and a pinned staging buffer.  Saving takes two passes like the reference
// (checksum for the header, then the links).
#include "lt_internal.hpp"

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <limits>
#include <sstream>
#include <string>
#include <sys/stat.h>

using lt::fail;
using lt::guarded;

namespace {

constexpr uint64_t ADLER_P = 4294967291ull;
constexpr size_t HEADER_BYTES = 4096;
constexpr int BS = 256;   // block size of the pack / unpack kernels
// the reference's program_version (LettuceGauge/defines.hpp:27), first line of bias files
const char* const BIAS_VERSION = "SU(3)_version_1.3";

bool file_exists(const char* path)
{
    struct stat st;
    return ::stat(path, &st) == 0;
}

// (B << 32) + A after A <- (A + 1) mod p, B <- B mod p (Checksum_Adler64::Finalize)
uint64_t adler_final(uint64_t A, uint64_t B) { return ((B % ADLER_P) << 32) + (A + 1) % ADLER_P; }

// lower-case hexadecimal without leading zeros (Checksum_Adler64::ReturnString)
std::string hex_of(uint64_t v)
{
    char buf[32];
    std::snprintf(buf, sizeof buf, "%llx", (unsigned long long)v);
    return buf;
}

struct DeviceBuf {
    void* p = nullptr;
    explicit DeviceBuf(size_t bytes)
    {
        if (bytes && hipMalloc(&p, bytes) != hipSuccess) throw lt::LtError{LT_ENOMEM, "BMW staging allocation failed"};
    }
    ~DeviceBuf() { (void)hipFree(p); }
};
struct PinnedBuf {
    void* p = nullptr;
    explicit PinnedBuf(size_t bytes)
    {
        if (bytes && hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess)
            throw lt::LtError{LT_ENOMEM, "BMW pinned staging allocation failed"};
    }
    ~PinnedBuf() { (void)hipHostFree(p); }
};
struct File {
    std::FILE* f = nullptr;
    File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

// BMW sites per slab: at most 256 MiB of link data
long long slab_sites(const lt::Geom& g, int K)
{
    const long long per_site = 4LL * K * 8;
    long long n = (256LL << 20) / per_site;
    return n < g.V ? n : (long long)g.V;
}

// Pack the field in BMW order; with `out` the big-endian words are written there.
// Returns the finalised checksum.
uint64_t pack_pass(lt_field f, int K, std::FILE* out)
{
    lt_ctx ctx = f->ctx;
    const long long V = f->g.V, slab = slab_sites(f->g, K);
    const size_t slab_bytes = (size_t)slab * 4 * K * sizeof(uint64_t);
    const long long nb = (slab + BS - 1) / BS;
    DeviceBuf d_out(out ? slab_bytes : 0), d_part(2 * nb * sizeof(uint64_t));
    PinnedBuf h_out(out ? slab_bytes : 0), h_part(2 * nb * sizeof(uint64_t));
    uint64_t A = 0, B = 0;
    for (long long n0 = 0; n0 < V; n0 += slab) {
        const long long cnt = (V - n0) < slab ? (V - n0) : slab;
        const long long blocks = (cnt + BS - 1) / BS;
        lt::k_bmw_pack(f, K, n0, cnt, (uint64_t*)d_out.p, (uint64_t*)d_part.p);
        LT_HIP_CHECK(hipMemcpyAsync(h_part.p, d_part.p, 2 * blocks * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                    ctx->stream));
        if (out)
            LT_HIP_CHECK(hipMemcpyAsync(h_out.p, d_out.p, (size_t)cnt * 4 * K * sizeof(uint64_t),
                                        hipMemcpyDeviceToHost, ctx->stream));
        LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        const uint64_t* hp = (const uint64_t*)h_part.p;
        for (long long b = 0; b < blocks; ++b) {
            A = (A + hp[2 * b]) % ADLER_P;
            B = (B + hp[2 * b + 1]) % ADLER_P;
        }
        if (out && std::fwrite(h_out.p, sizeof(uint64_t), (size_t)cnt * 4 * K, out) != (size_t)cnt * 4 * K)
            throw lt::LtError{LT_EIO, std::string("BMW write failed: ") + std::strerror(errno)};
    }
    return adler_final(A, B);
}

}  // namespace

extern "C" {

int lt_field_bmw_checksum(lt_field f, int full_rows, unsigned long long* checksum)
{
    if (!f || !checksum) return fail(LT_EINVAL, "lt_field_bmw_checksum: null argument");
    if (!f->su3) return fail(LT_ESTATE, "lt_field_bmw_checksum: the field must hold SU(3) links");
    return guarded([&] {
        LT_HIP_CHECK(hipSetDevice(f->ctx->device));
        *checksum = pack_pass(f, full_rows ? 18 : 12, nullptr);
        return LT_OK;
    });
}

int lt_field_save_bmw(lt_field f, const char* path, int full_rows, int overwrite, char* checksum_hex, int hex_len)
{
    if (!f || !path) return fail(LT_EINVAL, "lt_field_save_bmw: null argument");
    if (!f->su3) return fail(LT_ESTATE, "lt_field_save_bmw: the field must hold SU(3) links");
    if (file_exists(path) && !overwrite)
        return fail(LT_EIO, std::string("lt_field_save_bmw: ") + path + " exists (overwrite = 0)");
    return guarded([&] {
        LT_HIP_CHECK(hipSetDevice(f->ctx->device));
        const int K = full_rows ? 18 : 12;
        const std::string cs = hex_of(pack_pass(f, K, nullptr));
        // header: first line, a comment, then '\n' up to 4096 bytes (snprintf leaves
        // its terminating NUL after the comment, as in SaveConfigBMW)
        char header[HEADER_BYTES];
        std::memset(header, '\n', sizeof header);
        const int off = std::snprintf(header, sizeof header, "#%sBMW %d %d %d %d %s\n", full_rows ? "18" : "",
                                      f->g.L[1], f->g.L[2], f->g.L[3], f->g.L[0], cs.c_str());
        if (off < 0 || (size_t)off >= sizeof header) throw lt::LtError{LT_EIO, "BMW header overflow"};
        const int com = std::snprintf(header + off, sizeof header - off, "\nGenerated with Lettuce\n");
        if (com < 0 || (size_t)com >= sizeof header - off) throw lt::LtError{LT_EIO, "BMW header overflow"};
        File file(path, "wb");
        if (!file.f) throw lt::LtError{LT_EIO, std::string("cannot open ") + path + ": " + std::strerror(errno)};
        if (std::fwrite(header, 1, sizeof header, file.f) != sizeof header)
            throw lt::LtError{LT_EIO, std::string("BMW header write failed: ") + std::strerror(errno)};
        const uint64_t again = pack_pass(f, K, file.f);
        if (hex_of(again) != cs) throw lt::LtError{LT_ESTATE, "BMW checksum changed between passes"};
        if (std::fflush(file.f) != 0) throw lt::LtError{LT_EIO, std::string("BMW flush failed: ") + std::strerror(errno)};
        if (checksum_hex && hex_len > 0) std::snprintf(checksum_hex, (size_t)hex_len, "%s", cs.c_str());
        return LT_OK;
    });
}

int lt_field_load_bmw(lt_field f, const char* path, int full_rows, int* checksum_ok, double* max_deviation)
{
    if (!f || !path) return fail(LT_EINVAL, "lt_field_load_bmw: null argument");
    if (!f->su3) return fail(LT_ESTATE, "lt_field_load_bmw: the field must hold SU(3) links");
    return guarded([&] {
        LT_HIP_CHECK(hipSetDevice(f->ctx->device));
        lt_ctx ctx = f->ctx;
        File file(path, "rb");
        if (!file.f) throw lt::LtError{LT_EIO, std::string("cannot open ") + path + ": " + std::strerror(errno)};
        char header[HEADER_BYTES + 1];
        if (std::fread(header, 1, HEADER_BYTES, file.f) != HEADER_BYTES)
            throw lt::LtError{LT_EIO, "BMW header block shorter than 4096 bytes"};
        header[HEADER_BYTES] = '\0';
        int lx = 0, ly = 0, lz = 0, lt_ = 0;
        char cs_file[32] = {0};
        const int got = std::sscanf(header, full_rows ? "#18BMW %d %d %d %d %31s" : "#BMW %d %d %d %d %31s", &lx, &ly,
                                    &lz, &lt_, cs_file);
        if (got < 5) throw lt::LtError{LT_EIO, "header does not match the expected BMW format"};
        if (lx != f->g.L[1] || ly != f->g.L[2] || lz != f->g.L[3] || lt_ != f->g.L[0])
            throw lt::LtError{LT_ESTATE, "lattice shape in the file does not match the field"};
        const int K = full_rows ? 18 : 12;
        const long long V = f->g.V, slab = slab_sites(f->g, K);
        const size_t slab_bytes = (size_t)slab * 4 * K * sizeof(uint64_t);
        const long long nb = (slab + BS - 1) / BS;
        DeviceBuf d_in(slab_bytes), d_part(2 * nb * sizeof(uint64_t)), d_dev(nb * sizeof(double));
        PinnedBuf h_in(slab_bytes), h_part(2 * nb * sizeof(uint64_t)), h_dev(nb * sizeof(double));
        uint64_t A = 0, B = 0;
        double dev_row2 = 0.0;
        for (long long n0 = 0; n0 < V; n0 += slab) {
            const long long cnt = (V - n0) < slab ? (V - n0) : slab;
            const long long blocks = (cnt + BS - 1) / BS;
            const size_t words = (size_t)cnt * 4 * K;
            if (std::fread(h_in.p, sizeof(uint64_t), words, file.f) != words)
                throw lt::LtError{LT_EIO, "BMW file ends before the last link"};
            LT_HIP_CHECK(hipMemcpyAsync(d_in.p, h_in.p, words * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
            lt::k_bmw_unpack(f, K, n0, cnt, (const uint64_t*)d_in.p, (uint64_t*)d_part.p, (double*)d_dev.p);
            LT_HIP_CHECK(hipMemcpyAsync(h_part.p, d_part.p, 2 * blocks * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                        ctx->stream));
            LT_HIP_CHECK(hipMemcpyAsync(h_dev.p, d_dev.p, blocks * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            const uint64_t* hp = (const uint64_t*)h_part.p;
            const double* hd = (const double*)h_dev.p;
            for (long long b = 0; b < blocks; ++b) {
                A = (A + hp[2 * b]) % ADLER_P;
                B = (B + hp[2 * b + 1]) % ADLER_P;
                dev_row2 = hd[b] > dev_row2 ? hd[b] : dev_row2;
            }
        }
        if (checksum_ok) *checksum_ok = hex_of(adler_final(A, B)) == cs_file;
        // SU3::Tests::IsGroupElement (math/su3.hpp:568-595): |1 - U^+U| <= 1e-6; with
        // all three rows stored, also the stored row 2 against conj(row0 x row1)
        double dev = 0.0;
        long long nonfinite = 0;
        lt::k_check(f, &dev, &nonfinite);
        if (nonfinite) dev = std::numeric_limits<double>::infinity();
        if (dev_row2 > dev) dev = dev_row2;
        if (max_deviation) *max_deviation = dev;
        if (!(dev <= 1e-6)) throw lt::LtError{LT_ENUMERIC, "loaded links are not SU(3) within 1e-6"};
        return LT_OK;
    });
}

// ---- bias potential text files --------------------------------------------------

int lt_bias_info(lt_bias b, double* params, int* bins, unsigned long long* exceeded)
{
    if (!b) return fail(LT_EINVAL, "lt_bias_info: null bias");
    if (params) {
        params[0] = b->cv_min;
        params[1] = b->cv_max;
        params[2] = b->weight;
        params[3] = b->width;
        params[4] = b->wt_param;
        params[5] = b->threshold;
    }
    if (bins) *bins = b->bins;
    if (exceeded) *exceeded = b->exceeded;
    return LT_OK;
}

int lt_bias_save_parameters(lt_bias b, const char* path, int overwrite)
{
    if (!b || !path) return fail(LT_EINVAL, "lt_bias_save_parameters: null argument");
    return guarded([&] {
        std::ofstream ofs(path, overwrite ? (std::ios::out | std::ios::trunc) : (std::ios::out | std::ios::app));
        if (!ofs) throw lt::LtError{LT_EIO, std::string("cannot open ") + path};
        ofs << BIAS_VERSION << "\n"
            << "Metadynamics Potential" << "\n"
            << "CV_min: " << b->cv_min << "\n"
            << "CV_max: " << b->cv_max << "\n"
            << "bin_number: " << b->bins << "\n"
            << "weight: " << b->weight << "\n"
            << "width: " << b->width << "\n"
            << "well_tempered_parameter: " << b->wt_param << "\n"
            << "threshold_weight: " << b->threshold << "\n"
            << "END_METADYN_PARAMS" << "\n";
        if (!ofs) throw lt::LtError{LT_EIO, "bias parameter write failed"};
        return LT_OK;
    });
}

int lt_bias_save_potential(lt_bias b, const char* path, int overwrite)
{
    if (!b || !path) return fail(LT_EINVAL, "lt_bias_save_potential: null argument");
    return guarded([&] {
        std::ofstream ofs(path, overwrite ? (std::ios::out | std::ios::trunc) : (std::ios::out | std::ios::app));
        if (!ofs) throw lt::LtError{LT_EIO, std::string("cannot open ") + path};
        if (b->grid.empty()) {
            ofs << "\nexceeded_count: " << b->exceeded << "\n";
            return LT_OK;
        }
        for (size_t i = 0; i + 1 < b->grid.size(); ++i) ofs << b->grid[i] << ",";
        ofs << b->grid.back() << "\n"
            << "exceeded_count: " << b->exceeded << "\n";
        if (!ofs) throw lt::LtError{LT_EIO, "bias potential write failed"};
        return LT_OK;
    });
}

// Parameters from the lines before END_METADYN_PARAMS, the grid from the second to
// last line and the exceeded count from the last line (the newest snapshot of a
// file that SavePotential appended to).
int lt_bias_load_potential(lt_bias b, const char* path)
{
    if (!b || !path) return fail(LT_EINVAL, "lt_bias_load_potential: null argument");
    return guarded([&] {
        std::ifstream ifs(path);
        if (!ifs) throw lt::LtError{LT_EIO, std::string("cannot open ") + path};
        std::vector<std::string> lines;
        for (std::string line; std::getline(ifs, line);) lines.push_back(line);
        if (lines.empty() || lines[0] != BIAS_VERSION)
            throw lt::LtError{LT_EIO, std::string("bias file version differs from ") + BIAS_VERSION};
        double cv_min = b->cv_min, cv_max = b->cv_max, weight = b->weight, width = b->width, wt = b->wt_param,
               thr = b->threshold;
        int bins = b->bins;
        // the line, after leading white space, starts with the token (FindTokenEnd,
        // IO/string_manipulation.hpp:18-36)
        auto value_after = [](const std::string& line, const char* token, std::string& out) {
            const size_t first = line.find_first_not_of(" \f\n\r\t\v");
            if (first == std::string::npos || line.compare(first, std::strlen(token), token) != 0) return false;
            out = line.substr(first + std::strlen(token));
            return true;
        };
        for (size_t i = 1; i < lines.size(); ++i) {
            const std::string& l = lines[i];
            if (l.find("END_METADYN_PARAMS") != std::string::npos) break;
            std::string v;
            if (value_after(l, "CV_min: ", v)) cv_min = std::stod(v);
            else if (value_after(l, "CV_max: ", v)) cv_max = std::stod(v);
            else if (value_after(l, "bin_number: ", v)) bins = std::stoi(v);
            else if (value_after(l, "well_tempered_parameter: ", v)) wt = std::stod(v);
            else if (value_after(l, "threshold_weight: ", v)) thr = std::stod(v);
            else if (value_after(l, "weight: ", v)) weight = std::stod(v);
            else if (value_after(l, "width: ", v)) width = std::stod(v);
        }
        if (!(cv_min < cv_max)) throw lt::LtError{LT_EIO, "bias file: CV_min < CV_max required"};
        if (bins <= 0) throw lt::LtError{LT_EIO, "bias file: bin_number > 0 required"};
        // the file ends with "<grid>\nexceeded_count: n\n": its last two lines
        size_t last = lines.size();
        while (last > 0 && lines[last - 1].empty()) --last;
        if (last < 2) throw lt::LtError{LT_EIO, "bias file has no potential"};
        std::string grid_line = lines[last - 2], count_line = lines[last - 1];
        if (!grid_line.empty() && grid_line.back() == '\r') grid_line.pop_back();
        std::vector<double> grid;
        for (size_t p = 0; p < grid_line.size();) {
            const size_t q = grid_line.find(',', p);
            grid.push_back(std::stod(grid_line.substr(p, q == std::string::npos ? std::string::npos : q - p)));
            p = q == std::string::npos ? grid_line.size() : q + 1;
        }
        std::string cnt;
        if (!value_after(count_line, "exceeded_count: ", cnt)) throw lt::LtError{LT_EIO, "bias file: no exceeded_count"};
        if (grid.size() != (size_t)bins + 1) throw lt::LtError{LT_EIO, "bias file: grid size differs from bin_number + 1"};
        b->cv_min = cv_min;
        b->cv_max = cv_max;
        b->bins = bins;
        b->bin_width = (cv_max - cv_min) / bins;
        b->bin_width_inverse = 1.0 / b->bin_width;
        b->weight = weight;
        b->width = width;
        b->wt_param = wt;
        b->threshold = thr;
        b->grid = std::move(grid);
        b->exceeded = std::stoull(cnt);
        return LT_OK;
    });
}

}  // extern "C"
