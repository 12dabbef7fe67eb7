// Drives the xpcs::Corr facade (include/xpcs_corr.hpp) the way the
// reference's main.cpp:169-201 and :428-455 drive xpcs::Corr: rows of a
// SparseData filled by the caller, multiTau2 into zeroed [tau][W*H] float
// arrays, then normalizeG2s on those arrays.
//
//   corr_api_check <in.bin> <out.bin> [repeats]
//
// in.bin (little endian): int32 W, H, frames, dpl, average, has_flatfield;
// int32 dqmap[W*H], sqmap[W*H]; float64 flatfield[W*H] if has_flatfield;
// int64 nnz; int64 row_off[W*H+1]; int32 frame[nnz]; float32 value[nnz].
// out.bin: int32 T, Q, max_level; int32 level[T], tau[T]; float32 G2, IP, IF
// [T][W*H] (multiTau2 called `repeats` times into the same arrays); float64
// g2, stderr [T][Q]. Exit status 0, or 2 with the error on stderr.
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <vector>

#include "xpcs_corr.hpp"

namespace {

template <typename T>
void get(FILE* f, T* p, size_t n)
{
    if (n && fread(p, sizeof(T), n, f) != n) throw std::runtime_error("input file too short");
}

template <typename T>
void put(FILE* f, const T* p, size_t n)
{
    if (n && fwrite(p, sizeof(T), n, f) != n) throw std::runtime_error("cannot write output file");
}

int run(int argc, char** argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: corr_api_check <in.bin> <out.bin> [repeats]\n");
        return 1;
    }
    const int repeats = argc > 3 ? atoi(argv[3]) : 1;
    FILE* in = fopen(argv[1], "rb");
    if (!in) throw std::runtime_error("cannot open input file");
    int32_t hdr[6];
    get(in, hdr, 6);
    const int W = hdr[0], H = hdr[1], F = hdr[2], dpl = hdr[3], avg = hdr[4];
    const size_t npix = (size_t)W * H;
    std::vector<int32_t> dq(npix), sq(npix);
    std::vector<double> ff;
    get(in, dq.data(), npix);
    get(in, sq.data(), npix);
    if (hdr[5]) {
        ff.resize(npix);
        get(in, ff.data(), npix);
    }
    int64_t nnz = 0;
    get(in, &nnz, 1);
    std::vector<int64_t> off(npix + 1);
    std::vector<int32_t> frame((size_t)nnz);
    std::vector<float> val((size_t)nnz);
    get(in, off.data(), npix + 1);
    get(in, frame.data(), (size_t)nnz);
    get(in, val.data(), (size_t)nnz);
    fclose(in);

    // rows as a Filter leaves them
    xpcs::data_structure::SparseData data((int)npix);
    for (size_t i = 0; i < npix; i++) {
        if (off[i + 1] == off[i]) continue;
        xpcs::data_structure::Row* r = data.Pixel((int)i);
        for (int64_t k = off[i]; k < off[i + 1]; k++) {
            r->indxPtr.push_back(frame[(size_t)k]);
            r->valPtr.push_back(val[(size_t)k]);
        }
    }

    xpcs_mt_desc d{};
    d.width = W;
    d.height = H;
    d.frames = F;
    d.delays_per_level = dpl;
    d.stride = 1;
    d.average = avg;
    d.dqmap = dq.data();
    d.sqmap = sq.data();
    d.flatfield = ff.empty() ? nullptr : ff.data();
    xpcs::Corr::configure(d);

    const int maxLevel = xpcs::Corr::calculateLevelMax(F, dpl);
    const auto delays = xpcs::Corr::delaysPerLevel(F, dpl, maxLevel);
    const size_t T = delays.size();
    std::vector<float> G2(T * npix, 0.f), IP(T * npix, 0.f), IF(T * npix, 0.f);
    for (int r = 0; r < repeats; r++) xpcs::Corr::multiTau2(&data, G2.data(), IP.data(), IF.data());
    int taus = 0, qbins = 0;
    xpcs::Corr::lastShape(&taus, &qbins);
    if ((size_t)taus != T) throw std::runtime_error("delaysPerLevel and the plan disagree on T");
    std::vector<double> g2(T * (size_t)qbins), se(T * (size_t)qbins);
    // main.cpp:441-445: the accumulated arrays as pixels x taus matrices
    xpcs::Corr::normalizeG2s(G2.data(), IP.data(), IF.data(), (int)npix, (int)T, g2.data(), se.data());
    xpcs::Corr::release();

    std::vector<int32_t> level(T), tau(T);
    for (size_t i = 0; i < T; i++) {
        level[i] = std::get<0>(delays[i]);
        tau[i] = std::get<1>(delays[i]);
    }
    FILE* out = fopen(argv[2], "wb");
    if (!out) throw std::runtime_error("cannot open output file");
    const int32_t shape[3] = {(int32_t)T, qbins, maxLevel};
    put(out, shape, 3);
    put(out, level.data(), T);
    put(out, tau.data(), T);
    put(out, G2.data(), G2.size());
    put(out, IP.data(), IP.size());
    put(out, IF.data(), IF.size());
    put(out, g2.data(), g2.size());
    put(out, se.data(), se.size());
    fclose(out);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        fprintf(stderr, "corr_api_check: %s\n", e.what());
        return 2;
    }
}
