// szx_batch.cpp -- the batched launch: many fields of mixed size and type
// compressed (and decompressed) by one launch sequence each way.
//
// The planner lays the caller's workspace out once (field tables, per-field
// look-back words, bitmaps and medians, record indexes, result words), uploads
// the tables, and afterwards every compress or decompress only enqueues
// kernels: no allocation, no copy, no synchronization, so a caller can capture
// the sequence in a hipGraph.  Each field's stream is exactly the stream of a
// single-field compress at the same arguments.
//
// Launch sequence of a batched compress:
//   k_zero_batch      clear look-back words, totals, error flags (flattened)
//   k_batch_range     REL fields only: min/max of 2^20-value items
//   k_batch_range_final   per REL field: the bound, rounded as the reference
//   k_compress_batch  one workgroup per tile of the global tile list
//   k_assemble_batch  header, bitmap, constant medians (flattened)
// and of a batched decompress:
//   k_decode_header_batch, k_decode_batch (one workgroup per index tile of
//   every field, the streams located through the result words).
//
// The tile shape of k_compress_batch (values per lane, index tiles per
// workgroup tile) is chosen per batch by szx_amd_batch_tune, which times every
// shape on the batch itself; SZX_AMD_BATCH_SHAPE="lv,ipt" overrides it.  The
// tuner's choice persists in a small JSON file, one entry per batch class
// (GPU arch and CU count, blockSize, fp32 share of the bytes, log2 of the
// bytes and of the field count): new plans of a known class start with its
// shape and the tuner does not time them again.  SZX_AMD_TUNE_CACHE names
// the file ("off" disables it); by default it is szx_amd_tune.json beside
// libszx_amd.so.
#include "szx_amd.h"

#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "szx_batch.hpp"
#include "szx_internal.hpp"

using namespace szxamd;

namespace {

constexpr uint64_t ALIGN = 256;
inline uint64_t al(uint64_t x) { return (x + ALIGN - 1) & ~(ALIGN - 1); }

int width_of(int dataType)
{
    if (dataType == SZx_FLOAT) return 4;
    if (dataType == SZx_DOUBLE) return 8;
    return 0;
}

// Byte offsets of one field's pieces inside the workspace.
struct FieldSlot {
    CompressLayout L;
    uint64_t ws;      // single-field compress workspace (compress_layout)
    uint64_t index;   // record index
    uint64_t hdr;     // decode header (256 bytes)
    uint64_t nidx;    // index tiles
    uint64_t ntiles[3];
    uint64_t range_items;
    uint64_t zero_items;
    uint64_t asm_items;
    int w;
};

struct Layout {
    uint64_t bf, df, bases, res, status, part, fields, dmap, drec, cmap[3], bytes;
    uint64_t total[3] = {0, 0, 0}, nitems = 0, dtotal = 0, zitems = 0, aitems = 0;
    std::vector<FieldSlot> slot;
};

constexpr int NBASE = 7;   // tbase (ipt 1), tbase (ipt 2), rbase, dbase, zbase, abase, tbase (ipt 4)

// Slot of a tile shape's counts: ntiles[] / total[], and the base arrays.
int tile_slot(int ipt) { return ipt == 4 ? 2 : ipt - 1; }
int tbase_slot(int ipt) { return ipt == 4 ? 6 : ipt - 1; }

bool plan_layout(const szx_amd_field *f, int nf, int bs, Layout &Y)
{
    if (nf <= 0 || !f) { api_error("no fields"); return false; }
    if (bs < 4 || bs > 256 || (bs & (bs - 1))) { api_error("unsupported blockSize %d", bs); return false; }
    Y.slot.resize(nf);
    uint64_t off = 0;
    Y.bf = off;     off = al(off + nf * sizeof(BatchField));
    Y.df = off;     off = al(off + nf * sizeof(BatchDecodeField));
    Y.bases = off;  off = al(off + NBASE * (nf + 1) * sizeof(uint64_t));
    Y.res = off;    off = al(off + 4 * (uint64_t)nf * sizeof(uint64_t));
    Y.status = off; off = al(off + (uint64_t)nf * sizeof(int32_t));
    for (int i = 0; i < nf; i++) {
        FieldSlot &s = Y.slot[i];
        s.w = width_of(f[i].dataType);
        if (!s.w) { api_error("field %d: unsupported dataType %d", i, f[i].dataType); return false; }
        if (f[i].n == 0) { api_error("field %d: n == 0", i); return false; }
        if (f[i].errBoundMode != ABS && f[i].errBoundMode != REL) {
            api_error("field %d: errBoundMode must be ABS or REL", i);
            return false;
        }
        const uint64_t n = f[i].n;
        s.L = compress_layout(n, s.w, bs);
        s.nidx = s.L.nidx;
        s.ntiles[0] = s.nidx;
        s.ntiles[1] = (n + INDEX_PER_TILE * tile_values(s.w) - 1) / (INDEX_PER_TILE * tile_values(s.w));
        s.ntiles[2] = (n + INDEX_PER_TILE_WIDE * tile_values(s.w) - 1) / (INDEX_PER_TILE_WIDE * tile_values(s.w));
        s.range_items = f[i].errBoundMode == REL ? (n + RANGE_ITEM - 1) / RANGE_ITEM : 0;
        s.zero_items = ((s.L.ws_misc + 64 * sizeof(uint64_t)) / 8 + ZERO_ITEM - 1) / ZERO_ITEM;
        const uint64_t work = s.L.nbit > 4 * s.L.nb ? s.L.nbit : 4 * s.L.nb;
        uint64_t blocks = (work + 256 * ASSEMBLE_BYTES_PER_THREAD - 1) / (256 * ASSEMBLE_BYTES_PER_THREAD);
        s.asm_items = blocks < 1 ? 1 : blocks > ASSEMBLE_MAX_BLOCKS ? ASSEMBLE_MAX_BLOCKS : blocks;
        Y.total[0] += s.ntiles[0];
        Y.total[1] += s.ntiles[1];
        Y.total[2] += s.ntiles[2];
        Y.nitems += s.range_items;
        Y.dtotal += s.nidx;
        Y.zitems += s.zero_items;
        Y.aitems += s.asm_items;
    }
    Y.part = off;   off = al(off + 2 * (Y.nitems ? Y.nitems : 1) * sizeof(double));
    Y.fields = off;
    for (int i = 0; i < nf; i++) {
        FieldSlot &s = Y.slot[i];
        s.ws = off;     off = al(off + s.L.ws_bytes);
        s.index = off;  off = al(off + 8 * index_words(f[i].n, s.w, bs));
        s.hdr = off;    off += ALIGN;
    }
    Y.dmap = off;   off = al(off + 4 * (Y.dtotal ? Y.dtotal : 1));
    Y.drec = off;   off = al(off + batch_decode_records_bytes(Y.dtotal ? Y.dtotal : 1));
    for (int k = 0; k < 3; k++) {   // the field of every compressor tile, per tile size
        Y.cmap[k] = off;
        off = al(off + 4 * (Y.total[k] ? Y.total[k] : 1));
    }
    Y.bytes = off;
    return true;
}

}  // namespace

struct szx_amd_batch {
    int nf = 0, bs = 0, dev = 0;
    uint8_t *work = nullptr;
    Layout Y;
    std::vector<szx_amd_field> fields;
    int lv = 8, ipt = 2;
    bool can_decode = false;
    std::vector<uint64_t> h_res;   // host copy of the result words (per-field decode)
    bool h_res_valid = false;

    BatchField *d_bf() const { return reinterpret_cast<BatchField *>(work + Y.bf); }
    BatchDecodeField *d_df() const { return reinterpret_cast<BatchDecodeField *>(work + Y.df); }
    uint64_t *d_base(int k) const
    {
        return reinterpret_cast<uint64_t *>(work + Y.bases) + (uint64_t)k * (nf + 1);
    }
    uint64_t *d_res() const { return reinterpret_cast<uint64_t *>(work + Y.res); }
    uint32_t *d_dmap() const { return reinterpret_cast<uint32_t *>(work + Y.dmap); }
    uint32_t *d_cmap(int k) const { return reinterpret_cast<uint32_t *>(work + Y.cmap[k]); }
    int32_t *d_status() const { return reinterpret_cast<int32_t *>(work + Y.status); }
    uint64_t *d_index(int i) const { return reinterpret_cast<uint64_t *>(work + Y.slot[i].index); }
};

namespace {

bool hip_ok(hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    api_error("%s: %s", what, hipGetErrorString(e));
    return false;
}

// ---- tuner cache --------------------------------------------------------

struct TunedShape {
    int lv = 0, ipt = 0;
    float ms = 0.0f;
};

std::mutex g_tune_mu;

// Path of the cache file, "" when disabled.
std::string tune_cache_path()
{
    if (const char *e = getenv("SZX_AMD_TUNE_CACHE")) {
        if (!strcmp(e, "off") || !strcmp(e, "0") || !*e) return "";
        return e;
    }
    Dl_info info{};
    if (!dladdr(reinterpret_cast<void *>(&tune_cache_path), &info) || !info.dli_fname) return "";
    std::string so = info.dli_fname;
    const size_t slash = so.rfind('/');
    return (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/szx_amd_tune.json";
}

// Class of a batch: the shapes' relative speed depends on the device, the
// block size, the dtype mix and the batch's size.
std::string tune_key(const szx_amd_batch *b)
{
    hipDeviceProp_t prop{};
    char arch[64] = "unknown";
    int cus = 0;
    if (hipGetDeviceProperties(&prop, b->dev) == hipSuccess) {
        snprintf(arch, sizeof arch, "%.60s", prop.gcnArchName);
        for (char *c = arch; *c; c++)
            if (*c == ':' || *c == '"' || *c == ' ') *c = '\0';   // feature suffixes off
        cus = prop.multiProcessorCount;
    }
    uint64_t bytes = 0, bytes32 = 0;
    for (const szx_amd_field &f : b->fields) {
        const uint64_t nb = f.n * (f.dataType == SZx_FLOAT ? 4ull : 8ull);
        bytes += nb;
        if (f.dataType == SZx_FLOAT) bytes32 += nb;
    }
    const int share = bytes ? (int)((10 * bytes32 + bytes / 2) / bytes) : 0;
    int lb = 0, lf = 0;
    while ((2ull << lb) <= bytes) lb++;
    while ((2 << lf) <= b->nf) lf++;
    char key[160];
    snprintf(key, sizeof key, "%s/%dcu/bs%d/f32share%d/bytes2^%d/fields2^%d", arch, cus, b->bs, share, lb, lf);
    return key;
}

// Entries of the file: lines of the form  "key": {"lanes": L, "tiles": T, "ms": M}
std::map<std::string, TunedShape> tune_cache_read(const std::string &path)
{
    std::map<std::string, TunedShape> m;
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return m;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        char *q0 = strchr(line, '"');
        char *q1 = q0 ? strchr(q0 + 1, '"') : nullptr;
        char *br = q1 ? strchr(q1, '{') : nullptr;
        if (!br) continue;
        TunedShape t;
        if (sscanf(br, "{\"lanes\": %d, \"tiles\": %d, \"ms\": %f}", &t.lv, &t.ipt, &t.ms) == 3)
            m[std::string(q0 + 1, q1)] = t;
    }
    fclose(f);
    return m;
}

bool tune_cache_lookup(const szx_amd_batch *b, TunedShape &t)
{
    const std::string path = tune_cache_path();
    if (path.empty()) return false;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    const auto m = tune_cache_read(path);
    const auto it = m.find(tune_key(b));
    if (it == m.end()) return false;
    t = it->second;
    return true;
}

// Adds or replaces the batch's entry (whole file rewritten through a
// temporary and a rename); a file that cannot be written is left alone.
void tune_cache_store(const szx_amd_batch *b, const TunedShape &t)
{
    const std::string path = tune_cache_path();
    if (path.empty()) return;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto m = tune_cache_read(path);
    m[tune_key(b)] = t;
    const std::string tmp = path + ".tmp" + std::to_string((long)getpid());
    FILE *f = fopen(tmp.c_str(), "w");
    if (!f) return;
    fputs("{\n", f);
    size_t k = 0;
    for (const auto &e : m)
        fprintf(f, "  \"%s\": {\"lanes\": %d, \"tiles\": %d, \"ms\": %.6g}%s\n", e.first.c_str(), e.second.lv,
                e.second.ipt, e.second.ms, ++k < m.size() ? "," : "");
    fputs("}\n", f);
    const bool ok = fclose(f) == 0;
    if (!ok || rename(tmp.c_str(), path.c_str()) != 0) (void)remove(tmp.c_str());
}

bool shape_ok(int bs, int lv, int ipt)
{
    if (ipt == 4) return (lv == 8 || lv == 16) && bs >= 16 && bs <= 128;
    return (ipt == 1 || ipt == 2) && (lv == 4 || (lv == 8 && bs >= 8));
}

int enqueue_compress(szx_amd_batch *b, int mode, hipStream_t st)
{
    const Layout &Y = b->Y;
    b->h_res_valid = false;
    if (mode == SZX_AMD_BATCH_PER_FIELD) {
        // one single-field compress per field: the launches a caller without
        // the batched path would make
        for (int i = 0; i < b->nf; i++) {
            const szx_amd_field &f = b->fields[i];
            const FieldSlot &s = Y.slot[i];
            if (compress_launch(f.d_in, f.n, s.w, f.errBoundMode == REL, f.absErrBound, f.relBoundRatio, b->bs,
                                static_cast<uint8_t *>(f.d_out), b->work + s.ws, b->d_res() + 4 * i,
                                b->d_index(i), st))
                return SZx_NSCS;
        }
        return hip_ok(hipGetLastError(), "launch") ? SZx_SCES : SZx_NSCS;
    }
    if (mode != SZX_AMD_BATCH_FUSED) { api_error("unknown batch mode %d", mode); return SZx_NSCS; }
    batch_zero_launch(b->d_bf(), b->d_base(4), b->nf, Y.zitems, st);
    if (Y.nitems) batch_range_launch(b->d_bf(), b->d_base(2), b->nf, Y.nitems,
                                     reinterpret_cast<double *>(b->work + Y.part), st);
    timing_mark_begin(KK_COMPRESS, st);
    const int rc = batch_compress_launch(b->bs, b->lv, b->ipt, b->d_bf(), b->d_base(tbase_slot(b->ipt)),
                                         b->d_cmap(tile_slot(b->ipt)), Y.total[tile_slot(b->ipt)], st);
    timing_mark_end(KK_COMPRESS, st);
    if (rc) { api_error("unsupported tile shape"); return SZx_NSCS; }
    batch_assemble_launch(b->d_bf(), b->d_base(5), b->nf, Y.aitems, st);
    return hip_ok(hipGetLastError(), "launch") ? SZx_SCES : SZx_NSCS;
}

}  // namespace

extern "C" {

size_t szx_amd_batch_workspace(const szx_amd_field *fields, int nfields, int blockSize)
{
    Layout Y;
    return plan_layout(fields, nfields, blockSize, Y) ? Y.bytes : 0;
}

szx_amd_batch *szx_amd_batch_create(const szx_amd_field *fields, int nfields, int blockSize, void *d_work,
                                    size_t workBytes, void *hipStream)
{
    api_clear_error();
    Layout Y;
    if (!plan_layout(fields, nfields, blockSize, Y)) return nullptr;
    if (!d_work || ((uintptr_t)d_work & (ALIGN - 1))) { api_error("d_work must be 256-byte aligned"); return nullptr; }
    if (workBytes < Y.bytes) { api_error("workBytes < %llu", (unsigned long long)Y.bytes); return nullptr; }
    bool dec_all = true, dec_any = false;
    for (int i = 0; i < nfields; i++) {
        const szx_amd_field &f = fields[i];
        if (!f.d_in || !f.d_out) { api_error("field %d: NULL d_in or d_out", i); return nullptr; }
        if (((uintptr_t)f.d_in & 15) || ((uintptr_t)f.d_dec & 15)) {
            api_error("field %d: d_in and d_dec must be 16-byte aligned", i);
            return nullptr;
        }
        dec_all = dec_all && f.d_dec;
        dec_any = dec_any || f.d_dec;
    }
    if (dec_any && !dec_all) { api_error("give d_dec for every field or for none"); return nullptr; }

    szx_amd_batch *b = new szx_amd_batch;
    b->nf = nfields;
    b->bs = blockSize;
    b->work = static_cast<uint8_t *>(d_work);
    b->Y = Y;
    b->fields.assign(fields, fields + nfields);
    b->can_decode = dec_all;
    (void)hipGetDevice(&b->dev);
    TunedShape t;
    if (const char *e = getenv("SZX_AMD_BATCH_SHAPE")) {
        int lv = 0, ipt = 0;
        if (sscanf(e, "%d,%d", &lv, &ipt) == 2 && shape_ok(blockSize, lv, ipt)) {
            b->lv = lv;
            b->ipt = ipt;
        }
    } else if (tune_cache_lookup(b, t) && shape_ok(blockSize, t.lv, t.ipt)) {
        b->lv = t.lv;
        b->ipt = t.ipt;
    }
    if (!shape_ok(blockSize, b->lv, b->ipt)) b->lv = 4;

    // field tables
    std::vector<BatchField> bf(nfields);
    std::vector<BatchDecodeField> df(nfields);
    std::vector<uint64_t> base(NBASE * (nfields + 1), 0);
    uint64_t acc[NBASE] = {};
    for (int i = 0; i < nfields; i++) {
        const szx_amd_field &f = fields[i];
        const FieldSlot &s = Y.slot[i];
        uint8_t *ws = b->work + s.ws;
        unsigned long long *misc = reinterpret_cast<unsigned long long *>(ws + s.L.ws_misc);
        BatchField &B = bf[i];
        memset(&B, 0, sizeof B);
        CompressParams &p = B.p;
        p.x = f.d_in;
        p.n = f.n;
        p.eb = f.errBoundMode == ABS ? f.absErrBound : 0.0f;
        p.eb_dev = f.errBoundMode == REL ? reinterpret_cast<float *>(misc + 32) : nullptr;
        p.payload = static_cast<uint8_t *>(f.d_out) + s.L.region;
        p.bitmap = ws + s.L.ws_bitmap;
        p.cmed = reinterpret_cast<uint32_t *>(ws + s.L.ws_cmed);
        p.tiles = reinterpret_cast<unsigned long long *>(ws + s.L.ws_tiles);
        p.chunks = reinterpret_cast<ChunkState *>(ws + s.L.ws_chunks);
        p.cacc = reinterpret_cast<unsigned long long *>(ws + s.L.ws_cacc);
        p.totals = reinterpret_cast<uint64_t *>(misc + 16);
        p.err = reinterpret_cast<int *>(misc + 48);
        p.index = b->d_index(i);
        p.bidx = reinterpret_cast<uint16_t *>(p.index + 2 * (s.nidx + 1));
        p.nidx = s.nidx;
        p.nb = s.L.nb;
        p.ntiles = s.ntiles[1];
        p.nbit = s.L.nbit;
        AssembleParams &a = B.a;
        a.out = static_cast<uint8_t *>(f.d_out);
        a.region = s.L.region;
        a.bitmap = p.bitmap;
        a.cmed = p.cmed;
        a.totals = p.totals;
        a.nbit = s.L.nbit;
        a.bs = blockSize;
        a.eb_dev = p.eb_dev;
        a.eb = p.eb;
        a.result = b->d_res() + 4 * i;
        a.err = p.err;
        B.chunks = p.chunks;
        B.ntiles[0] = s.ntiles[0];
        B.ntiles[1] = s.ntiles[1];
        B.ntiles[2] = s.ntiles[2];
        B.zwords = (s.L.ws_misc + 64 * sizeof(uint64_t)) / 8;
        B.relv = f.relBoundRatio;
        B.rel = f.errBoundMode == REL;
        B.w = s.w;

        BatchDecodeField &D = df[i];
        memset(&D, 0, sizeof D);
        D.out = static_cast<const uint8_t *>(f.d_out);
        D.result = a.result;
        D.n = f.n;
        D.index = p.index;
        D.dst = f.d_dec;
        D.status = b->d_status() + i;
        D.h = reinterpret_cast<DecodeHeader *>(b->work + s.hdr);
        D.ntidx = s.nidx;
        D.w = s.w;

        const uint64_t cnt[NBASE] = {s.ntiles[0], s.ntiles[1], s.range_items, s.nidx,
                                     s.zero_items, s.asm_items, s.ntiles[2]};
        for (int k = 0; k < NBASE; k++) {
            base[k * (nfields + 1) + i] = acc[k];
            acc[k] += cnt[k];
        }
    }
    for (int k = 0; k < NBASE; k++) base[k * (nfields + 1) + nfields] = acc[k];
    // the field of every decode tile (one load per workgroup instead of a
    // search of the tile bases)
    std::vector<uint32_t> dmap(Y.dtotal ? Y.dtotal : 1, 0);
    for (int i = 0; i < nfields; i++)
        for (uint64_t t = 0; t < Y.slot[i].nidx; t++) dmap[base[3 * (nfields + 1) + i] + t] = (uint32_t)i;
    // the same for the compressor's tiles of each size (tile_slot order)
    std::vector<uint32_t> cmap[3];
    for (int k = 0; k < 3; k++) {
        const int bslot = k == 2 ? 6 : k;   // tbase_slot of ipt 1, 2, 4
        cmap[k].assign(Y.total[k] ? Y.total[k] : 1, 0);
        for (int i = 0; i < nfields; i++)
            for (uint64_t t = 0; t < Y.slot[i].ntiles[k]; t++)
                cmap[k][base[bslot * (nfields + 1) + i] + t] = (uint32_t)i;
    }

    hipStream_t st = static_cast<hipStream_t>(hipStream);
    const bool ok = hip_ok(hipMemcpyAsync(b->d_bf(), bf.data(), bf.size() * sizeof(BatchField),
                                          hipMemcpyHostToDevice, st), "H2D fields") &&
                    hip_ok(hipMemcpyAsync(b->d_df(), df.data(), df.size() * sizeof(BatchDecodeField),
                                          hipMemcpyHostToDevice, st), "H2D fields") &&
                    hip_ok(hipMemcpyAsync(b->d_base(0), base.data(), base.size() * sizeof(uint64_t),
                                          hipMemcpyHostToDevice, st), "H2D bases") &&
                    hip_ok(hipMemcpyAsync(b->d_dmap(), dmap.data(), dmap.size() * sizeof(uint32_t),
                                          hipMemcpyHostToDevice, st), "H2D tile map") &&
                    hip_ok(hipMemcpyAsync(b->d_cmap(0), cmap[0].data(), cmap[0].size() * sizeof(uint32_t),
                                          hipMemcpyHostToDevice, st), "H2D tile map") &&
                    hip_ok(hipMemcpyAsync(b->d_cmap(1), cmap[1].data(), cmap[1].size() * sizeof(uint32_t),
                                          hipMemcpyHostToDevice, st), "H2D tile map") &&
                    hip_ok(hipMemcpyAsync(b->d_cmap(2), cmap[2].data(), cmap[2].size() * sizeof(uint32_t),
                                          hipMemcpyHostToDevice, st), "H2D tile map") &&
                    hip_ok(hipStreamSynchronize(st), "batch setup");
    if (!ok) {
        delete b;
        return nullptr;
    }
    return b;
}

void szx_amd_batch_destroy(szx_amd_batch *b) { delete b; }

int szx_amd_batch_compress(szx_amd_batch *b, int mode, void *hipStream)
{
    api_clear_error();
    if (!b) { api_error("NULL batch"); return SZx_NSCS; }
    return enqueue_compress(b, mode, static_cast<hipStream_t>(hipStream));
}

uint64_t *szx_amd_batch_results_dev(szx_amd_batch *b) { return b ? b->d_res() : nullptr; }

int32_t *szx_amd_batch_status_dev(szx_amd_batch *b) { return b ? b->d_status() : nullptr; }

uint64_t *szx_amd_batch_index_dev(szx_amd_batch *b, int field)
{
    return b && field >= 0 && field < b->nf ? b->d_index(field) : nullptr;
}

int szx_amd_batch_results(szx_amd_batch *b, uint64_t *h_results, void *hipStream)
{
    api_clear_error();
    if (!b) { api_error("NULL batch"); return SZx_NSCS; }
    hipStream_t st = static_cast<hipStream_t>(hipStream);
    b->h_res.resize(4 * (size_t)b->nf);
    if (!hip_ok(hipMemcpyAsync(b->h_res.data(), b->d_res(), b->h_res.size() * sizeof(uint64_t),
                               hipMemcpyDeviceToHost, st), "D2H results") ||
        !hip_ok(hipStreamSynchronize(st), "batch results"))
        return SZx_NSCS;
    b->h_res_valid = true;
    if (h_results) memcpy(h_results, b->h_res.data(), b->h_res.size() * sizeof(uint64_t));
    return SZx_SCES;
}

int szx_amd_batch_decompress(szx_amd_batch *b, int flags, int mode, void *hipStream)
{
    api_clear_error();
    if (!b) { api_error("NULL batch"); return SZx_NSCS; }
    if (!b->can_decode) { api_error("the batch was created without d_dec"); return SZx_NSCS; }
    hipStream_t st = static_cast<hipStream_t>(hipStream);
    const int exact = flags & SZX_AMD_DECODE_EXACT;
    if (mode == SZX_AMD_BATCH_PER_FIELD) {
        // the single-field decompress per field; it takes the stream's
        // position and length from the host, like szx_amd_decompress_dev
        if (!b->h_res_valid) { api_error("call szx_amd_batch_results after the compress first"); return SZx_NSCS; }
        for (int i = 0; i < b->nf; i++) {
            const szx_amd_field &f = b->fields[i];
            const FieldSlot &s = b->Y.slot[i];
            const uint8_t *stream = static_cast<const uint8_t *>(f.d_out) + b->h_res[4 * i];
            if (decompress_launch(stream, b->h_res[4 * i + 1], f.n, s.w, b->bs, f.d_dec, b->work + s.hdr, ALIGN,
                                  b->d_index(i), exact, b->d_status() + i, st))
                return SZx_NSCS;
        }
        return hip_ok(hipGetLastError(), "launch") ? SZx_SCES : SZx_NSCS;
    }
    if (mode != SZX_AMD_BATCH_FUSED) { api_error("unknown batch mode %d", mode); return SZx_NSCS; }
    timing_mark_begin(KK_DECODE, st);
    const int rc = batch_decode_launch(b->bs, b->d_df(), b->d_base(3), b->d_dmap(), b->work + b->Y.drec, b->nf,
                                       b->Y.dtotal, exact, st);
    timing_mark_end(KK_DECODE, st);
    if (rc) { api_error("unsupported blockSize"); return SZx_NSCS; }
    return hip_ok(hipGetLastError(), "launch") ? SZx_SCES : SZx_NSCS;
}

int szx_amd_batch_set_shape(szx_amd_batch *b, int valuesPerLane, int tilesPerGroup)
{
    api_clear_error();
    if (!b || !shape_ok(b->bs, valuesPerLane, tilesPerGroup)) {
        api_error("unsupported tile shape (%d, %d)", valuesPerLane, tilesPerGroup);
        return SZx_NSCS;
    }
    b->lv = valuesPerLane;
    b->ipt = tilesPerGroup;
    return SZx_SCES;
}

int szx_amd_batch_tune(szx_amd_batch *b, int reps, int *valuesPerLane, int *tilesPerGroup, float *ms,
                       void *hipStream)
{
    api_clear_error();
    if (!b) { api_error("NULL batch"); return SZx_NSCS; }
    TunedShape cached;
    if (tune_cache_lookup(b, cached) && shape_ok(b->bs, cached.lv, cached.ipt)) {
        b->lv = cached.lv;
        b->ipt = cached.ipt;
        if (valuesPerLane) *valuesPerLane = cached.lv;
        if (tilesPerGroup) *tilesPerGroup = cached.ipt;
        if (ms) *ms = cached.ms;
        return SZx_SCES;
    }
    if (reps < 1) reps = 1;
    hipStream_t st = static_cast<hipStream_t>(hipStream);
    const int shapes[6][2] = {{16, 4}, {8, 4}, {8, 2}, {8, 1}, {4, 2}, {4, 1}};
    const int lv0 = b->lv, ipt0 = b->ipt;   // restored if timing fails
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!hip_ok(hipEventCreate(&e0), "event") || !hip_ok(hipEventCreate(&e1), "event")) {
        if (e0) (void)hipEventDestroy(e0);
        return SZx_NSCS;
    }
    float best = 0.0f;
    int blv = 0, bipt = 0, rc = SZx_SCES;
    for (const auto &sh : shapes) {
        if (!shape_ok(b->bs, sh[0], sh[1])) continue;
        b->lv = sh[0];
        b->ipt = sh[1];
        // one untimed pass (code object load, caches), then reps timed passes
        if ((rc = enqueue_compress(b, SZX_AMD_BATCH_FUSED, st)) != SZx_SCES) break;
        if (!hip_ok(hipEventRecord(e0, st), "event")) { rc = SZx_NSCS; break; }
        for (int r = 0; r < reps && rc == SZx_SCES; r++) rc = enqueue_compress(b, SZX_AMD_BATCH_FUSED, st);
        float t = 0.0f;
        if (rc != SZx_SCES || !hip_ok(hipEventRecord(e1, st), "event") ||
            !hip_ok(hipEventSynchronize(e1), "tune") || !hip_ok(hipEventElapsedTime(&t, e0, e1), "event")) {
            rc = SZx_NSCS;
            break;
        }
        t /= reps;
        if (!blv || t < best) {
            best = t;
            blv = sh[0];
            bipt = sh[1];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != SZx_SCES || !blv) {
        b->lv = lv0;
        b->ipt = ipt0;
        return rc != SZx_SCES ? rc : SZx_NSCS;
    }
    b->lv = blv;
    b->ipt = bipt;
    TunedShape t;
    t.lv = blv;
    t.ipt = bipt;
    t.ms = best;
    tune_cache_store(b, t);
    if (valuesPerLane) *valuesPerLane = blv;
    if (tilesPerGroup) *tilesPerGroup = bipt;
    if (ms) *ms = best;
    return SZx_SCES;
}

}  // extern "C"
