// Frame records of the packed detector formats (host side, no device work).
//
// UFXC and Rigaku files are flat arrays of packed words (u32 / u64); their
// readers decide which output frame every word belongs to.  That decision is
// sequential (a wrapping frame counter, a flush state machine), so it runs
// here over the words once; the words themselves go to the device unchanged
// and the transpose decodes pixel and value (k_transpose.hip, tr_decode).  A
// record is a run of consecutive words of one output frame: (first word, word
// count, output frame), the same triple the IMM path has per raw frame.
#include <algorithm>
#include <vector>

#include "internal.h"

namespace xpcs {

namespace {

void push_run(std::vector<PackedRecord>& out, int64_t b, int64_t e, int32_t frame)
{
    // records hold at most 2^31 words
    for (int64_t x = b; x < e; x += (int64_t)1 << 31)
        out.push_back({x, (uint32_t)std::min<int64_t>(e - x, (int64_t)1 << 31), frame});
}

}  // namespace

// Ufxc::Ufxc (src/xpcs/io/ufxc.cpp:59-103): the raw frame of a word is its
// 11-bit counter (w >> 21) unwrapped against the previous word (a jump below
// -2000 adds 2048, above +2000 subtracts 2048), relative to the first word;
// the words of a raw frame need not be contiguous.  The load loop reads raw
// frames from 0 -- Ufxc::SkipFrames does nothing (:173-175) and UFXC files
// have no dark frames -- in blocks of stride*avg (main.cpp:245-262): output
// frame f sums raw frames f*stride*avg + j*stride, j < avg.  Raw frames past
// the data are empty (Ufxc::NextFrames, :120-131).
xpcs_status ufxc_records(const uint32_t* w, int64_t n, const HostTables& h, std::vector<PackedRecord>& out)
{
    out.clear();
    if (n <= 0) return XPCS_OK;
    struct Run {
        int64_t ff, b, e;
    };
    std::vector<Run> runs;
    uint32_t value = w[0] >> 21;
    const int64_t f0 = value;
    int64_t bf = 0, cur = 0, start = 0;
    for (int64_t i = 1; i < n; i++) {
        const uint32_t c = w[i] >> 21;
        const int64_t diff = (int64_t)c - (int64_t)value;
        if (diff < -2000) bf += 2048;
        else if (diff > 2000) bf -= 2048;
        const int64_t ff = (int64_t)c + bf - f0;
        value = c;
        if (ff != cur) {
            runs.push_back({cur, start, i});
            cur = ff;
            start = i;
        }
    }
    runs.push_back({cur, start, n});
    std::stable_sort(runs.begin(), runs.end(), [](const Run& a, const Run& b) { return a.ff < b.ff; });
    const int64_t R = (int64_t)h.stride * h.avg;
    for (int64_t f = 0; f < h.F; f++)
        for (int64_t j = 0; j < h.avg; j++) {
            const int64_t raw = f * R + j * h.stride;
            auto it = std::lower_bound(runs.begin(), runs.end(), raw,
                                       [](const Run& r, int64_t v) { return r.ff < v; });
            for (; it != runs.end() && it->ff == raw; ++it) push_run(out, it->b, it->e, (int32_t)f);
        }
    return XPCS_OK;
}

// Rigaku::Rigaku (src/xpcs/io/rigaku.cpp:128-255): words of frames up to
// data_begin_todo - 1 are skipped; with stride > 1 so are words of frames not
// divisible by the stride (frame 0 excepted).  The pending output frame is
// flushed -- whether or not it holds a pixel -- when a word's frame differs
// from the previous flush's (avg 1) or passes the next expected frame (avg > 1,
// which then advances by stride*avg): output frames are counted in flushes, so
// a file frame without words makes no output frame.  Words of masked pixels
// are not pending pixels (the transpose drops them on the device; here they
// stay in their run).  Reading stops at `frames` output frames.  The last pending
// frame is kept only when it holds more than one distinct pixel (:233).
xpcs_status rigaku_records(const uint64_t* w, int64_t n, const HostTables& h, std::vector<PackedRecord>& out)
{
    out.clear();
    const int64_t W = h.W, H = h.H, npd = W * H;
    const int64_t start = h.frame_from, R = (int64_t)h.stride * h.avg;
    int64_t real = 0, next_expected = start + R, previous = start + 1;
    std::vector<PackedRecord> pending;
    // Only the frame field decides output frames, so the machine steps over
    // runs of words with one frame number; inside a run each word can cause at
    // most one flush, and once a word does not, no later word of the run does.
    // Pixels are checked by the transpose on the device (words of records only,
    // as the reference decodes only the words it keeps).
    auto flush_due = [&](int64_t frame) {
        return (h.avg > 1 && frame > next_expected) || (h.avg == 1 && frame != previous);
    };
    bool done = false;
    for (int64_t b = 0; b < n && !done;) {
        const uint64_t fw = w[b] >> 40;
        int64_t e = b + 1;
        while (e < n && (w[e] >> 40) == fw) e++;
        const int64_t frame = (int64_t)fw;
        int64_t i = b;
        b = e;
        if (frame <= start) continue;
        if (real >= h.F) break;
        if (h.stride > 1 && frame != 0 && frame % h.stride != 0) continue;
        while (i < e && flush_due(frame)) {
            // word i flushes the pending frame and opens frame real + 1
            out.insert(out.end(), pending.begin(), pending.end());
            pending.clear();
            previous = frame;
            real++;
            next_expected += R;
            if (real >= h.F) {   // the next word stops the reader; frame `frames` is never kept
                done = true;
                break;
            }
            push_run(pending, i, i + 1, (int32_t)real);
            i++;
        }
        if (done || i >= e) continue;
        // the rest of the run joins the pending frame (merged with a one-word run just before it)
        if (!pending.empty() && pending.back().word + (int64_t)pending.back().count == i &&
            pending.back().frame == (int32_t)real && pending.back().count < (1u << 31)) {
            const int64_t b0 = pending.back().word;
            pending.pop_back();
            push_run(pending, b0, e, (int32_t)real);
        } else {
            push_run(pending, i, e, (int32_t)real);
        }
    }
    if (!done && real < h.F && !pending.empty()) {
        // the last pending frame is kept with more than one distinct mapped pixel
        std::vector<int64_t> pix;
        for (const auto& r : pending)
            for (int64_t k = r.word; k < r.word + (int64_t)r.count; k++) {
                const uint32_t cm = (uint32_t)(w[k] >> 16) & 0xfffffu;
                const int64_t p = (int64_t)(cm % (uint32_t)H) * W + cm / (uint32_t)H;
                if (p < npd && h.pix2c[(size_t)p] >= 0) pix.push_back(p);
            }
        std::sort(pix.begin(), pix.end());
        if (std::unique(pix.begin(), pix.end()) - pix.begin() > 1) out.insert(out.end(), pending.begin(), pending.end());
    }
    return XPCS_OK;
}

}  // namespace xpcs

extern "C" xpcs_status xpcs_packed_records(const xpcs_mt_desc* desc, const void* words, size_t nbytes,
                                           int32_t format, int64_t* first_word, uint32_t* count,
                                           int32_t* frame, int64_t cap, int64_t* n)
{
    if (!desc || (!words && nbytes) || !n) return xpcs::fail(XPCS_ERR_INVALID_ARG, "packed_records: NULL argument");
    xpcs::HostTables h;
    xpcs_status st = xpcs::build_tables(*desc, h);
    if (st) return st;
    std::vector<xpcs::PackedRecord> rec;
    st = xpcs::packed_records(words, nbytes, format, h, rec);
    if (st) return st;
    *n = (int64_t)rec.size();
    for (int64_t i = 0; i < cap && i < (int64_t)rec.size(); i++) {
        if (first_word) first_word[i] = rec[(size_t)i].word;
        if (count) count[i] = rec[(size_t)i].count;
        if (frame) frame[i] = rec[(size_t)i].frame;
    }
    return XPCS_OK;
}

namespace xpcs {

xpcs_status packed_records(const void* words, size_t nbytes, int32_t format, const HostTables& h,
                           std::vector<PackedRecord>& rec)
{
    if (format == XPCS_FMT_UFXC) {
        if (nbytes % 4) return fail(XPCS_ERR_FORMAT, "ufxc: %zu bytes is not a whole number of 32-bit words", nbytes);
        return ufxc_records(static_cast<const uint32_t*>(words), (int64_t)(nbytes / 4), h, rec);
    }
    if (format == XPCS_FMT_RIGAKU) {
        if (nbytes % 8)
            return fail(XPCS_ERR_FORMAT, "rigaku: %zu bytes is not a whole number of 64-bit words", nbytes);
        return rigaku_records(static_cast<const uint64_t*>(words), (int64_t)(nbytes / 8), h, rec);
    }
    return fail(XPCS_ERR_INVALID_ARG, "packed words: unknown format %d", format);
}

}  // namespace xpcs
