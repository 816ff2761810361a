// stz_cli.cpp -- command-line driver with the reference program's grammar:
//
//   stz <error_bound> <raw_file> <dim_x> [dim_y dim_z] <-f|-d> [--chunks high low]
//       [--query-only] [--auto-query] [x0 y0 z0 x1 y1 z1]
//       [--rel | --abs] [--trial] [--out decoded.raw] [--save-stream file.stz]
//
// It mirrors main/run_typed of
//   F5 = src/zero_all_line_dynamic_random_access_optfast3_optcr_lb_fastdec5.cpp
// (argument grammar F5:1957-2070, raw I/O F5:1913-1949, report lines
// F5:2071-2086, F5:2314, F5:2642-2666, F5:2960, F5:3003-3016) on top of the
// C ABI: compress on the GPU, decompress, PSNR, and a random-access query
// checked against the crop of the full decode.  error_bound keeps F5's
// meaning (level bounds are mult_* x error_bound, STZ_EB_REF) unless --rel /
// --abs make it the maximum pointwise error.
#include <stz.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

namespace {

const char* kUsage =
    " <error_bound> <raw_file> <dim_x> [dim_y dim_z] <-f|-d> [--chunks high low] [--query-only] "
    "[--auto-query] [x0 y0 z0 x1 y1 z1] [--rel|--abs] [--trial] [--out file] [--save-stream file]";

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool read_file(const std::string& path, std::vector<char>& buf, size_t expect)
{
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) {
        std::cerr << "Failed to open file for reading: " << path << std::endl;
        return false;
    }
    const size_t n = (size_t)f.tellg();
    if (expect && n != expect) {
        std::cerr << "Input file size mismatch: expected " << expect << " bytes, got " << n << " bytes." << std::endl;
        return false;
    }
    buf.resize(n);
    f.seekg(0);
    f.read(buf.data(), (std::streamsize)n);
    return (bool)f;
}

bool write_file(const std::string& path, const void* p, size_t n)
{
    std::ofstream f(path, std::ios::binary);
    if (!f) {
        std::cerr << "Failed to open file for writing: " << path << std::endl;
        return false;
    }
    f.write((const char*)p, (std::streamsize)n);
    return (bool)f;
}

int fail(const char* what, int rc)
{
    std::cerr << what << " failed: " << stz_strerror(rc) << " (" << rc << ")" << std::endl;
    return 1;
}

template <typename T>
int run(int argc, char** argv, int type_arg)
{
    const double eb = std::atof(argv[1]);
    const std::string path = argv[2];
    const int fx = std::atoi(argv[3]);
    const int fy = type_arg == 4 ? fx : std::atoi(argv[4]);
    const int fz = type_arg == 4 ? fx : std::atoi(argv[5]);
    if (fx <= 0 || fy <= 0 || fz <= 0) {
        std::cerr << "dims must be positive, got " << fx << " x " << fy << " x " << fz << std::endl;
        return 1;
    }
    stz_params prm;
    stz_default_params(&prm);
    stz_eb_mode mode = STZ_EB_REF;
    bool full_validate = true, auto_query = false;
    std::string out_path, stream_path;
    std::vector<int> q;
    for (int a = type_arg + 1; a < argc; ++a) {
        const std::string t = argv[a];
        if (t == "--chunks") {
            if (a + 2 >= argc) {
                std::cerr << "--chunks requires two integers: <high_dim> <low_dim>" << std::endl;
                return 1;
            }
            prm.chunk_high = std::atoi(argv[a + 1]);
            prm.chunk_low = std::atoi(argv[a + 2]);
            if (prm.chunk_high <= 0 || prm.chunk_low <= 0) {
                std::cerr << "--chunks dims must be positive, got " << argv[a + 1] << " " << argv[a + 2] << std::endl;
                return 1;
            }
            a += 2;
        } else if (t == "--query-only" || t == "--no-full-validate") {
            full_validate = false;
            auto_query = true;
        } else if (t == "--auto-query") {
            auto_query = true;
        } else if (t == "--rel") {
            mode = STZ_EB_REL;
        } else if (t == "--abs") {
            mode = STZ_EB_ABS;
        } else if (t == "--trial") {
            prm.tuner = 2;
        } else if ((t == "--out" || t == "--save-stream") && a + 1 < argc) {
            (t == "--out" ? out_path : stream_path) = argv[++a];
        } else {
            char* end = nullptr;
            const long v = std::strtol(argv[a], &end, 10);
            if (*argv[a] == '\0' || *end != '\0') {
                std::cerr << "Invalid argument '" << argv[a]
                          << "'. Expected ROI integer, --query-only, or --auto-query." << std::endl;
                return 1;
            }
            q.push_back((int)v);
        }
    }
    if (prm.tuner == 2 && mode == STZ_EB_REF) {
        std::cerr << "--trial needs --rel or --abs (the bound must not depend on the configuration)" << std::endl;
        return 1;
    }
    if (!q.empty() && q.size() != 6) {
        std::cerr << "ROI query must provide exactly six integers: x0 y0 z0 x1 y1 z1" << std::endl;
        return 1;
    }
    if (q.empty()) q = {0, 0, 0, std::min(64, fx) - 1, std::min(64, fy) - 1, std::min(64, fz) - 1};  // F5:33

    const size_t n = (size_t)fx * fy * fz;
    std::vector<char> raw;
    if (!read_file(path, raw, n * sizeof(T))) return 1;
    const T* field = (const T*)raw.data();
    const stz_dtype dt = sizeof(T) == 4 ? STZ_F32 : STZ_F64;

    std::cout << "Input: " << path << ", dims: " << fx << " x " << fy << " x " << fz
              << ", type: " << (sizeof(T) == 4 ? "float" : "double") << std::endl;
    std::cout << "Random-access Huffman markers: high " << prm.chunk_high << "^3, low " << prm.chunk_low << "^3"
              << std::endl;
    if (!full_validate)
        std::cout << "Mode: query-only random access (skip full fine decode and PSNR validation)." << std::endl;

    stz_ctx* ctx = nullptr;
    int rc = stz_create(&ctx, 0);
    if (rc != STZ_OK) return fail("stz_create", rc);

    // compress: host field -> device payload (timed), then the zstd container
    const double g0 = now();
    stz_dstream* ds = nullptr;
    rc = stz_compress_device(ctx, dt, field, fx, fy, fz, eb, mode, &prm, &ds);
    if (rc != STZ_OK) return fail("compress", rc);
    const double g1 = now();
    void* stream = nullptr;
    size_t bytes = 0;
    rc = stz_dstream_download(ctx, ds, prm.zstd_level, &stream, &bytes);
    if (rc != STZ_OK) return fail("download", rc);
    const double g2 = now();
    std::cout << "CR: " << (double)(n * sizeof(T)) / (double)bytes << std::endl;
    stz_info info;
    stz_dstream_info(ds, &info);
    std::cout << "payload bytes before zstd: " << info.payload_bytes << ", stream bytes: " << bytes
              << ", level bounds base/low/high: " << info.eb_base << " / " << info.eb_low << " / " << info.eb_high
              << std::endl;
    if (!stream_path.empty() && !write_file(stream_path, stream, bytes)) return 1;

    // decompress from the container (upload + zstd + GPU decode), timed
    std::vector<T> dec;
    double d_dev = 0, d_glob = 0;
    if (full_validate) {
        dec.resize(n);
        const double t0 = now();
        stz_dstream* ds2 = nullptr;
        rc = stz_dstream_upload(ctx, stream, bytes, &ds2);
        if (rc != STZ_OK) return fail("upload", rc);
        const double t1 = now();
        rc = stz_dstream_decompress(ctx, ds2, dec.data());
        if (rc != STZ_OK) return fail("decompress", rc);
        const double t2 = now();
        d_dev = t2 - t1;
        d_glob = t2 - t0;
        stz_dstream_free(ds2);
        double mse = 0, lo = 1e300, hi = -1e300, maxe = 0;
        for (size_t i = 0; i < n; ++i) {
            const double o = (double)field[i], df = o - (double)dec[i];
            mse += df * df;
            lo = std::min(lo, o);
            hi = std::max(hi, o);
            maxe = std::max(maxe, std::fabs(df));
        }
        mse /= (double)n;
        std::cout << "Global PSNR: " << 20 * std::log10(hi - lo) - 10 * std::log10(mse) << std::endl;
        std::cout << "max abs error: " << maxe << " (bound " << info.eb_high << ")" << std::endl;
        if (!out_path.empty() && !write_file(out_path, dec.data(), n * sizeof(T))) return 1;
    }

    // random-access query on the resident stream (F5:2460-2666)
    if (auto_query || full_validate) {
        std::vector<T> roi((size_t)(std::abs(q[3] - q[0]) + 1) * (std::abs(q[4] - q[1]) + 1) * (std::abs(q[5] - q[2]) + 1));
        size_t od[3];
        const double t0 = now();
        rc = stz_dstream_decompress_roi(ctx, ds, q[0], q[1], q[2], q[3], q[4], q[5], roi.data(), od);
        if (rc != STZ_OK) return fail("ROI decompress", rc);
        const double t1 = now();
        // the normalised box is the output's extent
        const int x0 = std::max(0, std::min(q[0], q[3])), y0 = std::max(0, std::min(q[1], q[4])),
                  z0 = std::max(0, std::min(q[2], q[5]));
        const int bx = q[0] == q[3] ? std::min(std::max(q[0], 0), fx - 1) : x0;
        const int by = q[1] == q[4] ? std::min(std::max(q[1], 0), fy - 1) : y0;
        const int bz = q[2] == q[5] ? std::min(std::max(q[2], 0), fz - 1) : z0;
        std::cout << "Query [" << bx << "," << bx + (int)od[0] - 1 << "] x [" << by << "," << by + (int)od[1] - 1
                  << "] x [" << bz << "," << bz + (int)od[2] - 1 << "] "
                  << ((od[0] == 1 || od[1] == 1 || od[2] == 1) ? "(slice)" : "(box)") << std::endl;
        std::cout << "Auto query end-to-end decompress time: " << std::fixed << std::setprecision(5) << t1 - t0
                  << " sec" << std::endl;
        std::cout.unsetf(std::ios::floatfield);
        if (full_validate) {
            double md = 0;
            for (size_t z = 0; z < od[2]; ++z)
                for (size_t y = 0; y < od[1]; ++y)
                    for (size_t x = 0; x < od[0]; ++x) {
                        const size_t i = ((size_t)(bz + z) * fy + (by + y)) * fx + (bx + x);
                        md = std::max(md, std::fabs((double)roi[(z * od[1] + y) * od[0] + x] - (double)dec[i]));
                    }
            std::cout << "Auto query final ROI max abs diff: " << md << std::endl;
        }
    }

    std::cout << "compress time is: " << std::fixed << std::setprecision(5) << g1 - g0 << " sec" << std::endl;
    std::cout << "decompress time is: " << std::fixed << std::setprecision(5) << d_dev << " sec" << std::endl;
    std::cout << "global compress time is: " << std::fixed << std::setprecision(5) << g2 - g0 << " sec" << std::endl;
    std::cout << "global decompress time is: " << std::fixed << std::setprecision(5) << d_glob << " sec" << std::endl;
    stz_free(stream);
    stz_dstream_free(ds);
    stz_destroy(ctx);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    int type_arg = -1;
    if (argc >= 5 && (std::string(argv[4]) == "-f" || std::string(argv[4]) == "-d")) type_arg = 4;
    else if (argc >= 7 && (std::string(argv[6]) == "-f" || std::string(argv[6]) == "-d")) type_arg = 6;
    if (type_arg < 0) {
        std::cerr << "Usage: " << argv[0] << kUsage << std::endl;
        return 1;
    }
    return std::string(argv[type_arg]) == "-f" ? run<float>(argc, argv, type_arg) : run<double>(argc, argv, type_arg);
}
