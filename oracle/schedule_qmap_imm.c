/*
 * Oracle part 1: tau schedule, q-partition map, IMM indexing and frame
 * selection.  TEST INFRASTRUCTURE ONLY (see xpcs_oracle.h).
 */
#include "xpcs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* src/xpcs/corr.cpp:1156-1160 */
int xo_level_max(int frames, int dpl)
{
    if (frames < dpl * 2) return 0;
    return (int)(floor(log2((double)frames) - log2(1.0 + 1.0 / (double)dpl)) - log2((double)dpl));
}

/* src/xpcs/corr.cpp:1162-1164 */
int xo_delay_count(int dpl, int level) { return level == 0 ? (2 * dpl - 1) : dpl; }

/* src/xpcs/corr.cpp:1133-1154 */
int xo_delays_per_level(int frames, int dpl, int max_level, int32_t* level, int32_t* tau, int cap)
{
    int n = 0;
    int ll_dpl = 0;
    for (int i = 0; i <= max_level; i++) {
        int step = (int)pow(2.0, i);
        int dpll = i == 0 ? dpl * 2 : dpl;
        for (int j = 0; j < dpll; j++) {
            if ((ll_dpl + step + pow(2.0, i)) > frames) break;
            if (n < cap) { level[n] = i; tau[n] = ll_dpl + step; }
            n++;
            ll_dpl = ll_dpl + step;
        }
    }
    return n;
}

/* ---- q map ---------------------------------------------------------------- */
typedef struct { int32_t q, s, idx; } qsp_t;

static int cmp_qsp(const void* a, const void* b)
{
    const qsp_t* x = (const qsp_t*)a;
    const qsp_t* y = (const qsp_t*)b;
    if (x->q != y->q) return x->q < y->q ? -1 : 1;
    if (x->s != y->s) return x->s < y->s ? -1 : 1;
    return (x->idx > y->idx) - (x->idx < y->idx);
}

/*
 * src/xpcs/configuration.cpp:244-381.  m_mapping{q -> {s -> [pixels]}} is a
 * nested std::map, so iteration is q ascending, s ascending, pixels in raster
 * order.  The sbin->qbin dedupe (:307-345) keeps the (q, s) group with the most
 * pixels (first q on ties) and erases the others; the pixels it meant to move
 * are pushed into a discarded copy (:327-328), so they leave m_mapping.
 */
int xo_build_qmap(int width, int height, const int32_t* dqmap, const int32_t* sqmap,
                  int32_t* order, int32_t* group_q, int32_t* group_s,
                  int32_t* group_begin, int32_t* group_end, int32_t* q_first_group,
                  int32_t* pixels_per_bin, int group_cap, int q_cap, int s_cap,
                  int32_t* counts)
{
    int64_t npix_det = (int64_t)width * height;
    int64_t npix = 0;
    int32_t max_q = 0, max_s = 0;
    for (int64_t i = 0; i < npix_det; i++) {
        if (dqmap[i] < 1 || sqmap[i] < 1) continue;
        npix++;
        if (dqmap[i] > max_q) max_q = dqmap[i];
        if (sqmap[i] > max_s) max_s = sqmap[i];
    }
    qsp_t* v = (qsp_t*)malloc(sizeof(qsp_t) * (size_t)(npix ? npix : 1));
    int64_t k = 0;
    for (int64_t i = 0; i < npix_det; i++) {
        if (dqmap[i] < 1 || sqmap[i] < 1) continue;
        v[k].q = dqmap[i]; v[k].s = sqmap[i]; v[k].idx = (int32_t)i; k++;
    }
    qsort(v, (size_t)npix, sizeof(qsp_t), cmp_qsp);

    /* groups before dedupe */
    int64_t ng0 = 0;
    for (int64_t i = 0; i < npix; i++)
        if (i == 0 || v[i].q != v[i - 1].q || v[i].s != v[i - 1].s) ng0++;
    int32_t* gq = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ng0 + 1));
    int32_t* gs = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ng0 + 1));
    int64_t* gb = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ng0 + 1));
    int64_t* ge = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ng0 + 1));
    char* keep = (char*)malloc((size_t)(ng0 + 1));
    int64_t g = -1;
    for (int64_t i = 0; i < npix; i++) {
        if (i == 0 || v[i].q != v[i - 1].q || v[i].s != v[i - 1].s) {
            g++; gq[g] = v[i].q; gs[g] = v[i].s; gb[g] = i; keep[g] = 1;
        }
        ge[g] = i + 1;
    }
    /* dedupe: for each sbin seen under more than one q, keep the largest group */
    int64_t* best = (int64_t*)malloc(sizeof(int64_t) * (size_t)(max_s + 1));
    int64_t* nq_of_s = (int64_t*)calloc((size_t)(max_s + 1), sizeof(int64_t));
    for (int s = 0; s <= max_s; s++) best[s] = -1;
    for (int64_t j = 0; j < ng0; j++) {           /* groups are q-ascending */
        int s = gs[j];
        nq_of_s[s]++;
        if (best[s] < 0 || (ge[j] - gb[j]) > (ge[best[s]] - gb[best[s]])) best[s] = j;
    }
    for (int64_t j = 0; j < ng0; j++)
        if (nq_of_s[gs[j]] > 1 && best[gs[j]] != j) keep[j] = 0;

    /* emit order: kept groups, then dropped pixels ascending */
    int64_t pos = 0, ng = 0;
    for (int64_t j = 0; j < ng0; j++) {
        if (!keep[j]) continue;
        if (ng < group_cap) {
            group_q[ng] = gq[j]; group_s[ng] = gs[j];
            group_begin[ng] = (int32_t)pos; group_end[ng] = (int32_t)(pos + ge[j] - gb[j]);
        }
        for (int64_t i = gb[j]; i < ge[j]; i++) order[pos++] = v[i].idx;
        ng++;
    }
    int64_t n_mapped = pos;
    /* dropped pixels in raster order */
    {
        int64_t d0 = pos;
        for (int64_t j = 0; j < ng0; j++) {
            if (keep[j]) continue;
            for (int64_t i = gb[j]; i < ge[j]; i++) order[pos++] = v[i].idx;
        }
        /* sort dropped ids ascending (insertion into a simple qsort) */
        int64_t nd = pos - d0;
        if (nd > 1) {
            qsp_t* tmp = (qsp_t*)malloc(sizeof(qsp_t) * (size_t)nd);
            for (int64_t i = 0; i < nd; i++) { tmp[i].q = 0; tmp[i].s = 0; tmp[i].idx = order[d0 + i]; }
            qsort(tmp, (size_t)nd, sizeof(qsp_t), cmp_qsp);
            for (int64_t i = 0; i < nd; i++) order[d0 + i] = tmp[i].idx;
            free(tmp);
        }
    }
    /* first group per q (1-based q), nq_out = max q */
    if (q_first_group && max_q + 1 <= q_cap) {
        int64_t gi = 0;
        for (int q = 1; q <= max_q; q++) {
            while (gi < ng && gi < group_cap && group_q[gi] < q) gi++;
            q_first_group[q - 1] = (int32_t)gi;
        }
        q_first_group[max_q] = (int32_t)ng;
    }
    if (pixels_per_bin && max_s <= s_cap) {
        for (int s = 0; s < max_s; s++) pixels_per_bin[s] = 0;
        for (int64_t i = 0; i < npix_det; i++) {
            if (sqmap[i] < 1 || dqmap[i] < 1) continue;
            pixels_per_bin[sqmap[i] - 1]++;
        }
    }
    /* number of q keys of m_mapping */
    int64_t nkeys = 0;
    for (int64_t i = 0; i < npix; i++)
        if (i == 0 || v[i].q != v[i - 1].q) nkeys++;

    counts[0] = (int32_t)npix;
    counts[1] = (int32_t)n_mapped;
    counts[2] = (int32_t)ng;
    counts[3] = max_q;
    counts[4] = max_s;
    counts[5] = (int32_t)nkeys;
    free(v); free(gq); free(gs); free(gb); free(ge); free(keep); free(best); free(nq_of_s);
    return ng <= group_cap ? 0 : -1;
}

/* ---- IMM ------------------------------------------------------------------ */
static uint32_t rd_u32(const uint8_t* p) { uint32_t x; memcpy(&x, p, 4); return x; }
static int32_t rd_i32(const uint8_t* p) { int32_t x; memcpy(&x, p, 4); return x; }
static double rd_f64(const uint8_t* p) { double x; memcpy(&x, p, 8); return x; }

/*
 * Header layout of src/xpcs/io/imm.h:63-144 (compression @4, elapsed @128,
 * dlen @152, corecotick @620; 1024 bytes).  As in Imm::Imm
 * (src/xpcs/io/imm.cpp:59-68) the first header decides compression for the
 * whole file; NextFrames (:70-118) reads dlen*4 index bytes when compressed,
 * then dlen*2 value bytes.
 */
int64_t xo_imm_index(const uint8_t* bytes, int64_t nbytes, xo_imm_frame* frames, int64_t cap)
{
    if (nbytes < 1024) return 0;
    int compressed = rd_i32(bytes + 4) != 0;
    int64_t per = compressed ? 6 : 2;
    int64_t off = 0, n = 0;
    while (off + 1024 <= nbytes) {
        const uint8_t* h = bytes + off;
        uint32_t dlen = rd_u32(h + 152);
        int64_t pay = off + 1024;
        if (pay + (int64_t)dlen * per > nbytes) return -1;
        if (n < cap) {
            frames[n].payload_off = pay;
            frames[n].dlen = dlen;
            frames[n].compressed = compressed;
            frames[n].elapsed = rd_f64(h + 128);
            frames[n].corecotick = rd_i32(h + 620);
            frames[n].pad = 0;
        }
        n++;
        off = pay + (int64_t)dlen * per;
    }
    return n;
}

/* src/xpcs/main.cpp:222-268 */
int64_t xo_select_frames(int64_t n_raw, int darks, int frame_from, int frames,
                         int stride, int avg, int64_t* sel, int32_t* out_frame, int64_t cap)
{
    int64_t r = 0;
    if (darks > 0) r += darks;
    if (frame_from > 0 && r < frame_from) r = frame_from;
    int64_t R = (int64_t)stride * avg;
    int64_t n = 0;
    for (int64_t f = 0; f < frames; f++) {
        for (int64_t j = 0; j < avg; j++) {
            int64_t raw = r + f * R + j * stride;
            if (raw >= n_raw) return -1;
            if (n < cap) { sel[n] = raw; out_frame[n] = (int32_t)f; }
            n++;
        }
    }
    return n;
}
