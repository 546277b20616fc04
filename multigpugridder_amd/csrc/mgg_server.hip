// Request server: continuous batching of forward / back projection requests
// against CAS volumes kept resident in HBM.
//
//   * Volumes: mgg_server_add_volume builds the CAS volume on the device
//     (Vol->CAS chain) once; it stays resident until removed.
//   * Admission: every request's device memory (per batch) comes out of one
//     arena of budget_bytes.  A batch starts when a pipeline is free and the
//     arena can hold it; otherwise it waits for running batches to finish.
//   * Batching: forward requests are merged into one batch of up to
//     max_batch_directions directions, across requests, across request
//     boundaries (a large request spans several batches) and across resident
//     volumes of different N: each sub-chunk of a batch is one forward
//     projection launch over work items of every volume in it
//     (launch_forward_project_multi), followed by each volume's image chain.  A
//     back request owns its accumulator and runs as one batch (its directions
//     in launches of up to max_batch_directions, accumulated in order).
//   * Overlap: num_pipelines batches run at once, each on its own stream, so
//     the transfers of one batch overlap the kernels of another and small
//     batches share the device.  Inside a batch the work runs in sub-chunks of
//     at most kSub directions, and host transfers go on the pipeline's second
//     (copy) stream: a forward batch copies each sub-chunk's images out while
//     the next sub-chunk is projected, a back request uploads the next chunk
//     of images (double-buffered) while the current one is projected, so a
//     large request no longer pays its whole PCIe transfer after its kernels.
//   * Results are those of the plain gridder path for the same request: a
//     forward request gets forwardProject's images, a back request gets getVol
//     of the back projection of its images into a zeroed volume.  A request's
//     result does not depend on what else is in flight: forward directions
//     are independent, and a back request accumulates only its own images in
//     a fixed order.
#include "mgg_common.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace mgg;

namespace {

using Clock = std::chrono::steady_clock;
constexpr size_t kAlign = 256;
size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// First-fit sub-allocator over one device allocation.
class Arena {
public:
    Arena() = default;
    Arena(const Arena &) = delete;
    Arena &operator=(const Arena &) = delete;
    ~Arena()
    {
        if (base_) (void)hipFree(base_);
    }
    void init(size_t bytes)
    {
        bytes = bytes & ~(kAlign - 1);
        hipError_t e = hipMalloc(&base_, bytes);
        if (e != hipSuccess) {
            base_ = nullptr;
            raise(MGG_ERR_OUT_OF_MEMORY, "server budget: hipMalloc of %zu bytes failed: %s", bytes,
                  hipGetErrorString(e));
        }
        size_ = bytes;
        free_[0] = bytes;
    }
    void *alloc(size_t bytes)
    {
        bytes = align_up(std::max<size_t>(bytes, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > bytes) free_[off + bytes] = len - bytes;
            live_[off] = bytes;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return base_ + off;
        }
        return nullptr;
    }
    void release(void *p)
    {
        if (!p) return;
        const size_t off = (size_t)((char *)p - base_);
        auto it = live_.find(off);
        if (it == live_.end()) return;
        size_t len = it->second, start = off;
        live_.erase(it);
        used_ -= len;
        auto nx = free_.lower_bound(start);
        if (nx != free_.end() && nx->first == start + len) {
            len += nx->second;
            free_.erase(nx);
        }
        auto pv = free_.lower_bound(start);
        if (pv != free_.begin()) {
            --pv;
            if (pv->first + pv->second == start) {
                start = pv->first;
                len += pv->second;
                free_.erase(pv);
            }
        }
        free_[start] = len;
    }
    size_t size() const { return size_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }

private:
    char *base_ = nullptr;
    size_t size_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;            // offset -> length
    std::unordered_map<size_t, size_t> live_;  // offset -> length
};

enum class Mem { Pageable, Pinned, Device };

Mem classify(const void *p)
{
    hipPointerAttribute_t at;
    std::memset(&at, 0, sizeof at);
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();  // an unregistered host pointer is not an error here
        return Mem::Pageable;
    }
    if (at.type == hipMemoryTypeDevice) return Mem::Device;
    if (at.type == hipMemoryTypeHost) return Mem::Pinned;
    return Mem::Pageable;
}

struct Volume {
    int id = 0;
    int N = 0, M = 0, V = 0, pad = 0;
    float maskRadius = 0.f;
    float *cas = nullptr;        // arena
    unsigned *check = nullptr;   // after cas: launch_check_volume's flag for cas and the table
    size_t bytes = 0;
    int users = 0;         // requests submitted and not finished
};

struct Req {
    unsigned long long ticket = 0;
    mgg_request r{};
    Volume *vol = nullptr;
    Mem axesMem = Mem::Pageable, imgMem = Mem::Pageable, volMem = Mem::Pageable;
    int issued = 0;  // directions handed to batches
    int done = 0;    // directions whose batch finished
    int status = MGG_OK;
    std::string err;
    double tSubmit = 0, tStart = -1, tDone = -1;
    bool finished = false;
};

struct Segment {
    Req *req;
    int first;   // first direction of the request in this batch
    int count;
    int offset;  // position in the batch
};

// Directions per sub-chunk inside a batch (transfer / compute overlap).
constexpr int kSub = 1024;

struct Pipeline {
    hipStream_t stream = nullptr;
    hipStream_t copy = nullptr;  // host transfers of the batch
    hipEvent_t done = nullptr;
    hipEvent_t handoff[2] = {nullptr, nullptr};  // compute -> copy (sub-chunk ready / buffer free)
    hipEvent_t copied[2] = {nullptr, nullptr};   // copy -> compute (chunk uploaded)
    hipEvent_t copyTail = nullptr;              // everything queued on the copy stream
    bool busy = false;
    float *stageAxes = nullptr;  // pinned
    size_t stageAxesLen = 0;
};

struct Batch {
    int kind = MGG_REQ_FORWARD;
    Volume *vol = nullptr;
    int n = 0;
    std::vector<Segment> segs;
    std::vector<size_t> imgBase;  // forward: first float of each segment's images in dImgs
    void *mem = nullptr;  // one arena block, subdivided
    Pipeline *pipe = nullptr;
    float *dImgs = nullptr;  // forward: every segment's images (N^2 each); back: N^3 volume
    int failed = MGG_OK;
    std::string err;
    bool ready = false;  // everything of the batch is queued on its stream
};

// Sub-buffers carved out of one arena block.
struct Carver {
    char *p;
    template <typename T> T *take(size_t bytes)
    {
        T *r = (T *)p;
        p += align_up(bytes);
        return r;
    }
};

size_t herm2d_bytes(int M, int n) { return sizeof(float2) * (size_t)std::min(n, (int)kFftBatch) * M * (M / 2 + 1); }

// 2-D transforms of n images in groups of kFftBatch; the last partial group is
// split into batches of 64, 16, 4 and 1 so that few plan shapes exist.
template <typename F>
void fft_groups(int n, F &&f)
{
    static const int sizes[5] = {(int)kFftBatch, 64, 16, 4, 1};
    int a = 0;
    for (int si = 0; si < 5 && a < n; si++)
        while (n - a >= sizes[si]) {
            f(a, sizes[si]);
            a += sizes[si];
        }
}

}  // namespace

struct mgg_server {
    int device = 0;
    int maxBatch = 8192;
    int nPipes = 4;
    int maxQueued = 1 << 20;
    Clock::time_point t0 = Clock::now();
    Arena arena;
    DevBuf dKer;
    std::mutex pcMu;
    std::map<int, std::unique_ptr<DevBuf>> precomp;  // per M, guarded by pcMu
    hipStream_t adminStream = nullptr;
    std::mutex adminMu;  // one volume build at a time on adminStream (its FFT plans and work memory)
    std::vector<std::unique_ptr<Pipeline>> pipes;

    std::mutex mu;
    std::condition_variable cvSched, cvDone, cvComp;
    std::deque<Req *> queue;
    std::unordered_map<unsigned long long, std::unique_ptr<Req>> reqs;
    std::map<int, std::unique_ptr<Volume>> volumes;
    std::list<std::unique_ptr<Batch>> inflight;
    unsigned long long nextTicket = 1;
    int nextVolume = 1;
    bool stop = false;
    bool hold = false;  // mgg_server_hold: queue requests, start no batch
    bool waitingForMemory = false;
    // stats
    unsigned long long submitted = 0, completed = 0, failedReqs = 0, batches = 0, batchDirs = 0,
                       mergedBatches = 0;
    size_t residentBytes = 0;
    std::thread scheduler, completer;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    ProjGeom geom(const Volume &v) const
    {
        ProjGeom g;
        g.V = v.V;
        g.M = v.M;
        g.maskRadius = v.maskRadius;
        g.ker = dKer.as<float>();
        g.K = 501;
        g.kerHWidth = 2.f;
        return g;
    }

    const float *precomp_for(int M)
    {
        std::lock_guard<std::mutex> lk(pcMu);
        auto &b = precomp[M];
        if (!b) {
            std::vector<double> pc;
            precomp_host(pc, M, 2);
            std::vector<float> pf(pc.begin(), pc.end());
            auto nb = std::make_unique<DevBuf>();
            nb->reserve(sizeof(float) * M);
            HIP_TRY(hipMemcpy(nb->ptr, pf.data(), sizeof(float) * M, hipMemcpyHostToDevice));
            b = std::move(nb);
        }
        return b->as<float>();
    }

    // ---- device memory of one batch -------------------------------------
    int sub_size() const { return std::min(maxBatch, kSub); }
    size_t fp_bytes(const Volume &v, int D) const
    {
        const int S = std::min(D, sub_size());
        return align_up(sizeof(float) * 9 * (size_t)D) + align_up(sizeof(float) * (size_t)S * v.M * v.M) +
               align_up(herm2d_bytes(v.M, S)) + align_up(sizeof(float) * (size_t)D * v.N * v.N);
    }
    // A forward batch whose segments may be on different volumes.
    size_t fp_bytes(const std::vector<Segment> &segs, int D) const
    {
        const int S = std::min(D, sub_size());
        int Mmax = 0;
        size_t imgs = 0;
        for (const Segment &sg : segs) {
            Mmax = std::max(Mmax, sg.req->vol->M);
            imgs += (size_t)sg.count * sg.req->vol->N * sg.req->vol->N;
        }
        return align_up(sizeof(float) * 9 * (size_t)D) + align_up(sizeof(float) * (size_t)S * Mmax * Mmax) +
               align_up(herm2d_bytes(Mmax, S)) + align_up(sizeof(float) * imgs);
    }
    size_t bp_bytes(const Volume &v, int n) const
    {
        const int C = std::min(n, sub_size());
        const size_t M = v.M;
        return align_up(sizeof(float) * 9 * (size_t)n) + 2 * align_up(sizeof(float) * (size_t)C * v.N * v.N) +
               align_up(sizeof(float) * (size_t)C * M * M) + align_up(herm2d_bytes(v.M, C)) +
               align_up(sizeof(float) * (size_t)v.V * v.V * v.V) + align_up(sizeof(float) * M * M * M) +
               align_up(sizeof(float2) * M * M * (M / 2 + 1)) + align_up(sizeof(float) * (size_t)v.N * v.N * v.N);
    }

    // ---- pinned staging of direction arrays -------------------------------
    float *stage_axes(Pipeline &p, size_t floats)
    {
        if (p.stageAxesLen < floats) {
            if (p.stageAxes) HIP_TRY(hipHostFree(p.stageAxes));
            p.stageAxes = nullptr;
            p.stageAxesLen = 0;
            const size_t want = std::max<size_t>(floats, 9 * (size_t)maxBatch);
            HIP_TRY(hipHostMalloc((void **)&p.stageAxes, sizeof(float) * want, hipHostMallocDefault));
            p.stageAxesLen = want;
        }
        return p.stageAxes;
    }

    void upload_axes(Pipeline &p, float *dAxes, const std::vector<Segment> &segs, int total)
    {
        float *st = stage_axes(p, 9 * (size_t)total);
        bool staged = false;
        for (const Segment &s : segs) {
            const float *src = s.req->r.axes + (size_t)9 * s.first;
            if (s.req->axesMem == Mem::Device) {
                HIP_TRY(hipMemcpyAsync(dAxes + (size_t)9 * s.offset, src, sizeof(float) * 9 * s.count,
                                       hipMemcpyDeviceToDevice, p.stream));
            } else {
                std::memcpy(st + (size_t)9 * s.offset, src, sizeof(float) * 9 * s.count);
                staged = true;
            }
        }
        if (staged) {
            // The whole range is copied; device-sourced segments are rewritten
            // after it in stream order.
            HIP_TRY(hipMemcpyAsync(dAxes, st, sizeof(float) * 9 * (size_t)total, hipMemcpyHostToDevice, p.stream));
            for (const Segment &s : segs)
                if (s.req->axesMem == Mem::Device)
                    HIP_TRY(hipMemcpyAsync(dAxes + (size_t)9 * s.offset, s.req->r.axes + (size_t)9 * s.first,
                                           sizeof(float) * 9 * s.count, hipMemcpyDeviceToDevice, p.stream));
        }
    }

    // ---- issue ------------------------------------------------------------
    // The copy stream's queued work joins the compute stream (the batch's
    // completion event is recorded after it).
    void join_copy(Pipeline &p)
    {
        HIP_TRY(hipEventRecord(p.copyTail, p.copy));
        HIP_TRY(hipStreamWaitEvent(p.stream, p.copyTail, 0));
    }

    void issue_forward(Batch &b)
    {
        Pipeline &p = *b.pipe;
        const int D = b.n, S = std::min(D, sub_size());
        int Mmax = 0;
        for (const Segment &sg : b.segs) Mmax = std::max(Mmax, sg.req->vol->M);
        Carver c{(char *)b.mem};
        float *dAxes = c.take<float>(sizeof(float) * 9 * (size_t)D);
        float *dCas = c.take<float>(sizeof(float) * (size_t)S * Mmax * Mmax);
        float2 *dHerm = c.take<float2>(herm2d_bytes(Mmax, S));
        b.imgBase.clear();
        size_t imgFloats = 0;
        for (const Segment &sg : b.segs) {
            b.imgBase.push_back(imgFloats);
            imgFloats += (size_t)sg.count * sg.req->vol->N * sg.req->vol->N;
        }
        float *dImgs = c.take<float>(sizeof(float) * imgFloats);
        b.dImgs = dImgs;
        upload_axes(p, dAxes, b.segs, D);
        int k = 0;
        for (int s0 = 0; s0 < D; s0 += S, k++) {
            const int cnt = std::min(S, D - s0);
            // The pieces of this sub-chunk: each segment's overlap with it,
            // their CAS images packed one after the other in dCas.
            struct Piece { int seg, lo, hi; size_t cas; };
            std::vector<Piece> pieces;
            size_t casOff = 0;
            for (int i = 0; i < (int)b.segs.size(); i++) {
                const Segment &sg = b.segs[i];
                const int lo = std::max(sg.offset, s0), hi = std::min(sg.offset + sg.count, s0 + cnt);
                if (lo >= hi) continue;
                pieces.push_back({i, lo, hi, casOff});
                const size_t M = (size_t)sg.req->vol->M;
                casOff += (size_t)(hi - lo) * M * M;
            }
            // One forward projection launch for up to kFpMaxSegments pieces,
            // whatever their volumes.
            for (size_t p0 = 0; p0 < pieces.size(); p0 += kFpMaxSegments) {
                FPMultiArgs ma;
                ma.nSegs = 0;
                unsigned work = 0;
                for (size_t q = p0; q < pieces.size() && q < p0 + kFpMaxSegments; q++) {
                    const Piece &pc = pieces[q];
                    const Volume &v = *b.segs[pc.seg].req->vol;
                    work += fp_segment(ma.seg[ma.nSegs++], v.cas, geom(v), dAxes + (size_t)9 * pc.lo, pc.hi - pc.lo,
                                       dCas + pc.cas, work, v.check);
                }
                ma.totalWork = work;
                check_launch(launch_forward_project_multi(ma, p.stream), "forward_project_multi");
            }
            // Each piece's inverse image chain (CASImgsToImgs), groups of kFftBatch images.
            for (const Piece &pc : pieces) {
                const Segment &sg = b.segs[pc.seg];
                const int M = sg.req->vol->M, N = sg.req->vol->N, n = pc.hi - pc.lo;
                const size_t mm = (size_t)M * M, nn = (size_t)N * N, hh = (size_t)M * (M / 2 + 1);
                const float scale = 1.0f / ((float)M * (float)M);
                float *cas = dCas + pc.cas;
                float *out = dImgs + b.imgBase[pc.seg] + (size_t)(pc.lo - sg.offset) * nn;
                for (int a0 = 0; a0 < n; a0 += (int)kFftBatch) {
                    const int gN = std::min((int)kFftBatch, n - a0);
                    if (fused_chain_supported(M, N)) {
                        check_launch(launch_fused_cas_to_images(cas + a0 * mm, dHerm, out + a0 * nn, M, N, gN, scale,
                                                                p.stream),
                                     "fused_cas_to_images");
                        continue;
                    }
                    check_launch(launch_cas_to_herm2d(cas + a0 * mm, dHerm, M, gN, p.stream), "cas_to_herm2d");
                    fft_groups(gN, [&](int a, int cn) { fft_exec(1, M, cn, dHerm + a * hh, cas + (a0 + a) * mm, p.stream); });
                    check_launch(launch_real2d_to_imgs(cas + a0 * mm, out + a0 * nn, M, N, gN, scale, p.stream),
                                 "real2d_to_imgs");
                }
            }
            // This sub-chunk's images go out on the copy stream while the next
            // sub-chunk is projected.
            HIP_TRY(hipEventRecord(p.handoff[k & 1], p.stream));
            HIP_TRY(hipStreamWaitEvent(p.copy, p.handoff[k & 1], 0));
            for (const Piece &pc : pieces) {
                const Segment &sg = b.segs[pc.seg];
                if (sg.req->imgMem == Mem::Pageable) continue;  // the completion thread copies it
                const size_t nn = (size_t)sg.req->vol->N * sg.req->vol->N;
                float *dst = sg.req->r.images + (size_t)(sg.first + pc.lo - sg.offset) * nn;
                const float *src = dImgs + b.imgBase[pc.seg] + (size_t)(pc.lo - sg.offset) * nn;
                HIP_TRY(hipMemcpyAsync(dst, src, sizeof(float) * (size_t)(pc.hi - pc.lo) * nn,
                                       sg.req->imgMem == Mem::Device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                       p.copy));
            }
        }
        join_copy(p);
        HIP_TRY(hipEventRecord(p.done, p.stream));
    }

    void issue_back(Batch &b)
    {
        const Volume &v = *b.vol;
        Pipeline &p = *b.pipe;
        Req &r = *b.segs[0].req;
        const int n = b.n, M = v.M, N = v.N, C = std::min(n, sub_size());
        const size_t mm = (size_t)M * M, nn = (size_t)N * N, hh = (size_t)M * (M / 2 + 1);
        Carver c{(char *)b.mem};
        float *dAxes = c.take<float>(sizeof(float) * 9 * (size_t)n);
        float *dImgBuf[2];
        dImgBuf[0] = c.take<float>(sizeof(float) * (size_t)C * nn);
        dImgBuf[1] = c.take<float>(sizeof(float) * (size_t)C * nn);
        float *dCas = c.take<float>(sizeof(float) * (size_t)C * mm);
        float2 *dHerm = c.take<float2>(herm2d_bytes(M, C));
        float *dAcc = c.take<float>(sizeof(float) * (size_t)v.V * v.V * v.V);
        float *dReal3 = c.take<float>(sizeof(float) * mm * M);
        float2 *dHerm3 = c.take<float2>(sizeof(float2) * mm * (M / 2 + 1));
        float *dVol = c.take<float>(sizeof(float) * nn * N);
        b.dImgs = dVol;
        upload_axes(p, dAxes, b.segs, n);
        const ProjGeom g = geom(v);
        check_launch(launch_fill(dAcc, 0.f, (size_t)v.V * v.V * v.V, p.stream), "fill");
        int k = 0;
        for (int c0 = 0; c0 < n; c0 += C, k++) {
            const int cnt = std::min(C, n - c0);
            const float *src = r.r.images + (size_t)c0 * nn;
            const float *imgs = src;
            if (r.imgMem != Mem::Device) {
                // Upload into the buffer the chunk before last used, once the
                // compute stream is done with it; double-buffered so the next
                // chunk's upload overlaps this chunk's kernels.  (Pageable
                // sources are staged by the runtime.)
                float *buf = dImgBuf[k & 1];
                if (k >= 2) HIP_TRY(hipStreamWaitEvent(p.copy, p.handoff[k & 1], 0));
                HIP_TRY(hipMemcpyAsync(buf, src, sizeof(float) * (size_t)cnt * nn, hipMemcpyHostToDevice, p.copy));
                HIP_TRY(hipEventRecord(p.copied[k & 1], p.copy));
                HIP_TRY(hipStreamWaitEvent(p.stream, p.copied[k & 1], 0));
                imgs = buf;
            }
            // forward image chain (ImgsToCASImgs), groups of kFftBatch images
            for (int a0 = 0; a0 < cnt; a0 += (int)kFftBatch) {
                const int gN = std::min((int)kFftBatch, cnt - a0);
                if (fused_chain_supported(M, N)) {
                    check_launch(launch_fused_images_to_cas(imgs + a0 * nn, dHerm, dCas + a0 * mm, M, N, gN, 1.0f, 0,
                                                            p.stream),
                                 "fused_images_to_cas");
                    continue;
                }
                check_launch(launch_imgs_to_real2d(imgs + a0 * nn, dCas + a0 * mm, M, N, gN, p.stream), "imgs_to_real2d");
                fft_groups(gN, [&](int a, int kk) { fft_exec(0, M, kk, dCas + (a0 + a) * mm, dHerm + a * hh, p.stream); });
                check_launch(launch_herm2d_to_cas(dHerm, dCas + a0 * mm, M, gN, p.stream), "herm2d_to_cas");
            }
            if (r.imgMem != Mem::Device) HIP_TRY(hipEventRecord(p.handoff[k & 1], p.stream));  // buffer free
            check_launch(back_project_launch(dAcc, nullptr, g, dAxes + (size_t)9 * c0, cnt, dCas, p.stream),
                         "back_project");
        }
        chain_cas_to_volume(dAcc, M, v.pad, dVol, N, precomp_for(M), 1, dReal3, dHerm3, p.stream);
        const size_t vbytes = sizeof(float) * nn * N;
        if (r.volMem == Mem::Device)
            HIP_TRY(hipMemcpyAsync(r.r.volume, dVol, vbytes, hipMemcpyDeviceToDevice, p.stream));
        else if (r.volMem == Mem::Pinned)
            HIP_TRY(hipMemcpyAsync(r.r.volume, dVol, vbytes, hipMemcpyDeviceToHost, p.stream));
        join_copy(p);
        HIP_TRY(hipEventRecord(p.done, p.stream));
    }

    // Host copies into pageable result arrays, after the batch finished.
    void finish_host_copies(Batch &b)
    {
        if (b.kind == MGG_REQ_FORWARD) {
            for (size_t i = 0; i < b.segs.size(); i++) {
                const Segment &s = b.segs[i];
                const size_t sn = (size_t)s.req->vol->N * s.req->vol->N;
                if (s.req->imgMem == Mem::Pageable)
                    HIP_TRY(hipMemcpy(s.req->r.images + (size_t)s.first * sn, b.dImgs + b.imgBase[i],
                                      sizeof(float) * (size_t)s.count * sn, hipMemcpyDeviceToHost));
            }
        } else {
            const Volume &v = *b.vol;  // a back request's own volume
            Req &r = *b.segs[0].req;
            if (r.volMem == Mem::Pageable)
                HIP_TRY(hipMemcpy(r.r.volume, b.dImgs, sizeof(float) * v.N * v.N * v.N, hipMemcpyDeviceToHost));
        }
    }

    // ---- scheduling ---------------------------------------------------------
    Pipeline *free_pipe()
    {
        for (auto &p : pipes)
            if (!p->busy) return p.get();
        return nullptr;
    }

    void fail_req(Req &r, int code, const std::string &msg)
    {
        if (r.finished) return;
        if (r.status == MGG_OK) {
            r.status = code;
            r.err = msg;
        }
        if (r.done >= r.issued && r.issued >= r.r.n_axes) finish_req(r);
    }

    void finish_req(Req &r)
    {
        if (r.finished) return;
        r.finished = true;
        r.tDone = now();
        r.vol->users--;
        if (r.status == MGG_OK)
            completed++;
        else
            failedReqs++;
        cvDone.notify_all();
    }

    // Forms the next batch from the queue head (lock held).  Returns nullptr
    // when the head batch does not fit in the arena now.
    std::unique_ptr<Batch> form_batch()
    {
        Req *h = queue.front();
        auto b = std::make_unique<Batch>();
        b->kind = h->r.kind;
        b->vol = h->vol;
        if (h->r.kind == MGG_REQ_BACK) {
            b->n = h->r.n_axes;
            b->segs.push_back({h, 0, h->r.n_axes, 0});
            b->mem = arena.alloc(bp_bytes(*h->vol, h->r.n_axes));
        } else {
            // Forward requests of every volume, in queue order.
            int D = 0;
            for (Req *r : queue) {
                if (D >= maxBatch) break;
                if (r->r.kind != MGG_REQ_FORWARD) continue;
                const int take = std::min(r->r.n_axes - r->issued, maxBatch - D);
                if (take <= 0) continue;
                b->segs.push_back({r, r->issued, take, D});
                D += take;
            }
            b->n = D;
            b->mem = arena.alloc(fp_bytes(b->segs, D));
        }
        if (!b->mem) return nullptr;
        const double t = now();
        std::vector<Req *> reqsInBatch;
        for (Segment &s : b->segs) {
            if (s.req->tStart < 0) s.req->tStart = t;
            s.req->issued += s.count;
            reqsInBatch.push_back(s.req);
        }
        queue.erase(std::remove_if(queue.begin(), queue.end(), [](Req *r) { return r->issued >= r->r.n_axes; }),
                    queue.end());
        batches++;
        batchDirs += (unsigned long long)b->n;
        if (b->segs.size() > 1) mergedBatches++;
        return b;
    }

    void scheduler_loop()
    {
        HIP_TRY(hipSetDevice(device));
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cvSched.wait(lk, [&] {
                return (stop && queue.empty()) || (!queue.empty() && !hold && free_pipe() && !waitingForMemory);
            });
            if (queue.empty()) {
                if (stop) return;
                continue;
            }
            Pipeline *p = free_pipe();
            std::unique_ptr<Batch> b = form_batch();
            if (!b) {
                if (inflight.empty()) {
                    // Nothing will free memory: the head cannot run at all.  Its
                    // unissued directions are abandoned (every issued batch has
                    // finished, since nothing is in flight).
                    Req *h = queue.front();
                    queue.pop_front();
                    h->issued = h->r.n_axes;
                    h->done = h->r.n_axes;
                    fail_req(*h, MGG_ERR_OUT_OF_MEMORY,
                             "request does not fit in the server's device-memory budget");
                } else {
                    waitingForMemory = true;  // cleared by the completion thread
                }
                continue;
            }
            b->pipe = p;
            p->busy = true;
            Batch *raw = b.get();
            inflight.push_back(std::move(b));
            lk.unlock();
            int code = MGG_OK;
            std::string msg;
            try {
                if (raw->kind == MGG_REQ_FORWARD)
                    issue_forward(*raw);
                else
                    issue_back(*raw);
            } catch (const Failure &f) {
                code = f.code;
                msg = g_last_error;
            } catch (const std::exception &e) {
                code = MGG_ERR_HIP;
                msg = e.what();
            }
            if (code != MGG_OK) {
                // Nothing more is queued for the batch; the event covers what
                // was queued before the failure on both streams, so its memory
                // is released only after that work.
                if (hipEventRecord(raw->pipe->copyTail, raw->pipe->copy) == hipSuccess)
                    (void)hipStreamWaitEvent(raw->pipe->stream, raw->pipe->copyTail, 0);
                (void)hipEventRecord(raw->pipe->done, raw->pipe->stream);
            }
            lk.lock();
            if (code != MGG_OK) {
                raw->failed = code;
                raw->err = msg;
            }
            raw->ready = true;
            cvComp.notify_all();
        }
    }

    void complete_batch(std::unique_ptr<Batch> b, int code, const std::string &msg)
    {
        // lock held
        for (Segment &s : b->segs) {
            s.req->done += s.count;
            if (code != MGG_OK) {
                fail_req(*s.req, code, msg);
            } else if (s.req->done >= s.req->r.n_axes) {
                finish_req(*s.req);
            }
            if (!s.req->finished && s.req->status != MGG_OK && s.req->done >= s.req->issued &&
                s.req->issued >= s.req->r.n_axes)
                finish_req(*s.req);
        }
        arena.release(b->mem);
        b->pipe->busy = false;
        waitingForMemory = false;
        cvSched.notify_all();
    }

    void completion_loop()
    {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cvComp.wait(lk, [&] { return stop || !inflight.empty(); });
            if (inflight.empty()) {
                if (stop && queue.empty()) return;
                continue;
            }
            bool any = false;
            for (auto it = inflight.begin(); it != inflight.end();) {
                Batch &b = **it;
                if (!b.ready) {
                    ++it;
                    continue;
                }
                hipError_t q = hipEventQuery(b.pipe->done);
                if (q == hipErrorNotReady) {
                    ++it;
                    continue;
                }
                std::unique_ptr<Batch> owned = std::move(*it);
                it = inflight.erase(it);
                any = true;
                int code = owned->failed;
                std::string msg = owned->err;
                if (code == MGG_OK && q != hipSuccess) {
                    code = MGG_ERR_HIP;
                    msg = std::string("batch failed on the device: ") + hipGetErrorString(q);
                }
                if (code == MGG_OK) {
                    lk.unlock();
                    try {
                        finish_host_copies(*owned);
                    } catch (const Failure &f) {
                        code = f.code;
                        msg = g_last_error;
                    }
                    lk.lock();
                }
                complete_batch(std::move(owned), code, msg);
            }
            if (!any) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lk.lock();
            }
        }
    }

    void start(const mgg_server_config &cfg)
    {
        device = cfg.device;
        if (cfg.max_batch_directions > 0) maxBatch = cfg.max_batch_directions;
        if (cfg.num_pipelines > 0) nPipes = cfg.num_pipelines;
        if (cfg.max_queued > 0) maxQueued = cfg.max_queued;
        DeviceGuard dg(device);
        size_t freeB = 0, totalB = 0;
        HIP_TRY(hipMemGetInfo(&freeB, &totalB));
        size_t budget = cfg.budget_bytes;
        if (budget == 0) budget = std::min<size_t>(freeB / 2, (size_t)32 << 30);
        if (budget > freeB) raise(MGG_ERR_OUT_OF_MEMORY, "budget %zu bytes exceeds free device memory %zu", budget, freeB);
        arena.init(budget);
        std::vector<double> w;
        kb_table_host(w, 501, 2);
        std::vector<float> wf(w.begin(), w.end());
        dKer.reserve(sizeof(float) * wf.size());
        HIP_TRY(hipMemcpy(dKer.ptr, wf.data(), sizeof(float) * wf.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipStreamCreateWithFlags(&adminStream, hipStreamNonBlocking));
        for (int i = 0; i < nPipes; i++) {
            auto p = std::make_unique<Pipeline>();
            HIP_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
            HIP_TRY(hipStreamCreateWithFlags(&p->copy, hipStreamNonBlocking));  // default priority
            HIP_TRY(hipEventCreateWithFlags(&p->done, hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&p->copyTail, hipEventDisableTiming));
            for (int e = 0; e < 2; e++) {
                HIP_TRY(hipEventCreateWithFlags(&p->handoff[e], hipEventDisableTiming));
                HIP_TRY(hipEventCreateWithFlags(&p->copied[e], hipEventDisableTiming));
            }
            pipes.push_back(std::move(p));
        }
        scheduler = std::thread([this] {
            try {
                scheduler_loop();
            } catch (...) {
            }
        });
        completer = std::thread([this] { completion_loop(); });
    }

    void shutdown()
    {
        {
            std::unique_lock<std::mutex> lk(mu);
            // every submitted request runs to completion first
            cvDone.wait(lk, [&] { return queue.empty() && inflight.empty(); });
            stop = true;
        }
        cvSched.notify_all();
        cvComp.notify_all();
        if (scheduler.joinable()) scheduler.join();
        if (completer.joinable()) completer.join();
        DeviceGuard dg(device);
        for (auto &p : pipes) {
            if (p->stream) {
                (void)hipStreamSynchronize(p->stream);
                fft_drop_stream(p->stream);
                (void)hipStreamDestroy(p->stream);
            }
            if (p->copy) {
                (void)hipStreamSynchronize(p->copy);
                (void)hipStreamDestroy(p->copy);
            }
            if (p->done) (void)hipEventDestroy(p->done);
            if (p->copyTail) (void)hipEventDestroy(p->copyTail);
            for (int e = 0; e < 2; e++) {
                if (p->handoff[e]) (void)hipEventDestroy(p->handoff[e]);
                if (p->copied[e]) (void)hipEventDestroy(p->copied[e]);
            }
            if (p->stageAxes) (void)hipHostFree(p->stageAxes);
        }
        if (adminStream) {
            (void)hipStreamSynchronize(adminStream);
            fft_drop_stream(adminStream);
            (void)hipStreamDestroy(adminStream);
        }
    }
};

namespace {
void need_ptr(const void *p, const char *what)
{
    if (!p) raise(MGG_ERR_INVALID_ARGUMENT, "%s is NULL", what);
}
}  // namespace

extern "C" {

int mgg_server_create(const mgg_server_config *cfg, mgg_server **out)
{
    return guarded([&] {
        need_ptr(out, "out");
        *out = nullptr;
        mgg_server_config c{};
        if (cfg) c = *cfg;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) raise(MGG_ERR_DEVICE, "no GPU visible");
        if (c.device < 0 || c.device >= ndev) raise(MGG_ERR_DEVICE, "GPU id %d out of range [0, %d)", c.device, ndev);
        if (c.max_batch_directions < 0 || c.num_pipelines < 0 || c.max_queued < 0)
            raise(MGG_ERR_INVALID_ARGUMENT, "negative server option");
        auto s = std::make_unique<mgg_server>();
        s->start(c);
        *out = s.release();
    });
}

int mgg_server_destroy(mgg_server *s)
{
    return guarded([&] {
        if (!s) return;
        s->shutdown();
        DeviceGuard dg(s->device);
        delete s;
    });
}

int mgg_server_add_volume(mgg_server *s, const float *volume, int volume_size, float interp_factor,
                          int extra_padding, int *volume_id)
{
    return guarded([&] {
        need_ptr(s, "server");
        need_ptr(volume, "volume");
        need_ptr(volume_id, "volume_id");
        if (volume_size <= 0) raise(MGG_ERR_INVALID_ARGUMENT, "volume size must be positive");
        if (!(interp_factor >= 1.f)) raise(MGG_ERR_INVALID_ARGUMENT, "interpFactor must be >= 1");
        if (extra_padding < 0) raise(MGG_ERR_INVALID_ARGUMENT, "extraPadding must be >= 0");
        auto v = std::make_unique<Volume>();
        v->N = volume_size;
        v->M = (int)((float)volume_size * interp_factor);  // gpuGridder.cpp:73-74
        v->pad = extra_padding;
        v->V = v->M + 2 * extra_padding;
        if (v->M % 2 || (v->M - v->N) % 2)
            raise(MGG_ERR_INVALID_ARGUMENT, "N*interpFactor = %d must be even and differ from N by an even number", v->M);
        v->maskRadius = ((float)volume_size * interp_factor) / 2 - 1;  // AbstractGridder.cpp:13
        DeviceGuard dg(s->device);
        const size_t M = v->M, N = v->N, V3 = (size_t)v->V * v->V * v->V;
        v->bytes = align_up(sizeof(float) * V3) + align_up(sizeof(unsigned));
        const size_t tmp = align_up(sizeof(float) * N * N * N) + align_up(sizeof(float) * M * M * M) +
                           align_up(sizeof(float2) * M * M * (M / 2 + 1));
        void *casMem = nullptr, *tmpMem = nullptr;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            casMem = s->arena.alloc(v->bytes);
            tmpMem = casMem ? s->arena.alloc(tmp) : nullptr;
            if (!casMem || !tmpMem) {
                s->arena.release(casMem);
                raise(MGG_ERR_OUT_OF_MEMORY,
                      "volume N=%d needs %zu bytes resident (+%zu while it is built); budget %zu, in use %zu",
                      v->N, v->bytes, tmp, s->arena.size(), s->arena.used());
            }
        }
        int code = MGG_OK;
        try {
            std::lock_guard<std::mutex> ak(s->adminMu);
            Carver c{(char *)tmpMem};
            float *dVol = c.take<float>(sizeof(float) * N * N * N);
            float *dReal3 = c.take<float>(sizeof(float) * M * M * M);
            float2 *dHerm3 = c.take<float2>(sizeof(float2) * M * M * (M / 2 + 1));
            HIP_TRY(hipMemcpyAsync(dVol, volume, sizeof(float) * N * N * N, hipMemcpyHostToDevice, s->adminStream));
            chain_volume_to_cas(dVol, v->N, (float *)casMem, v->M, v->pad, s->precomp_for(v->M), dReal3, dHerm3,
                                s->adminStream);
            // The volume is immutable while resident: one check serves every
            // forward batch on it.
            unsigned *check = (unsigned *)((char *)casMem + align_up(sizeof(float) * V3));
            check_launch(launch_check_volume(casMem, DT_F32, V3, s->dKer.as<float>(), 501, check, s->adminStream),
                         "check_volume");
            HIP_TRY(hipStreamSynchronize(s->adminStream));
        } catch (const Failure &f) {
            code = f.code;
        }
        std::lock_guard<std::mutex> lk(s->mu);
        s->arena.release(tmpMem);
        if (code != MGG_OK) {
            s->arena.release(casMem);
            throw Failure{code};
        }
        v->cas = (float *)casMem;
        v->check = (unsigned *)((char *)casMem + align_up(sizeof(float) * V3));
        v->id = s->nextVolume++;
        s->residentBytes += v->bytes;
        *volume_id = v->id;
        s->volumes[v->id] = std::move(v);
    });
}

int mgg_server_remove_volume(mgg_server *s, int volume_id)
{
    return guarded([&] {
        need_ptr(s, "server");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->volumes.find(volume_id);
        if (it == s->volumes.end()) raise(MGG_ERR_INVALID_ARGUMENT, "no resident volume %d", volume_id);
        Volume *v = it->second.get();
        s->cvDone.wait(lk, [&] { return v->users == 0; });
        s->arena.release(v->cas);
        s->residentBytes -= v->bytes;
        s->volumes.erase(it);
        s->waitingForMemory = false;
        s->cvSched.notify_all();
    });
}

int mgg_server_submit(mgg_server *s, const mgg_request *req, unsigned long long *ticket)
{
    return guarded([&] {
        need_ptr(s, "server");
        need_ptr(req, "request");
        need_ptr(ticket, "ticket");
        *ticket = 0;
        if (req->kind != MGG_REQ_FORWARD && req->kind != MGG_REQ_BACK)
            raise(MGG_ERR_INVALID_ARGUMENT, "request kind %d is neither forward nor back", req->kind);
        if (req->n_axes <= 0) raise(MGG_ERR_INVALID_ARGUMENT, "request needs at least one direction");
        need_ptr(req->axes, "request axes");
        need_ptr(req->images, "request images");
        if (req->kind == MGG_REQ_BACK) need_ptr(req->volume, "request volume");
        auto r = std::make_unique<Req>();
        r->r = *req;
        {
            DeviceGuard dg(s->device);
            r->axesMem = classify(req->axes);
            r->imgMem = classify(req->images);
            r->volMem = req->kind == MGG_REQ_BACK ? classify(req->volume) : Mem::Pageable;
        }
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) raise(MGG_ERR_INVALID_ARGUMENT, "server is shutting down");
        auto it = s->volumes.find(req->volume_id);
        if (it == s->volumes.end()) raise(MGG_ERR_INVALID_ARGUMENT, "no resident volume %d", req->volume_id);
        Volume *v = it->second.get();
        const size_t need = req->kind == MGG_REQ_BACK ? s->bp_bytes(*v, req->n_axes)
                                                      : s->fp_bytes(*v, std::min(req->n_axes, s->maxBatch));
        const size_t room = s->arena.size() - s->residentBytes;
        if (need > room)
            raise(MGG_ERR_OUT_OF_MEMORY, "request needs %zu bytes of device memory; the budget leaves %zu beside the resident volumes",
                  need, room);
        if ((int)s->queue.size() >= s->maxQueued)
            raise(MGG_ERR_BUSY, "request queue is full (%d waiting)", (int)s->queue.size());
        r->vol = v;
        v->users++;
        r->ticket = s->nextTicket++;
        r->tSubmit = s->now();
        *ticket = r->ticket;
        s->queue.push_back(r.get());
        s->reqs[r->ticket] = std::move(r);
        s->submitted++;
        s->cvSched.notify_all();
    });
}

int mgg_server_wait(mgg_server *s, unsigned long long ticket, double timeout_s, mgg_request_info *info)
{
    int code = MGG_OK;
    std::string msg;
    int rc = guarded([&] {
        need_ptr(s, "server");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) raise(MGG_ERR_INVALID_ARGUMENT, "unknown ticket %llu", ticket);
        Req *r = it->second.get();
        auto ready = [&] { return r->finished; };
        if (timeout_s < 0) {
            s->cvDone.wait(lk, ready);
        } else if (!s->cvDone.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
            raise(MGG_ERR_TIMEOUT, "request %llu not finished after %.3f s", ticket, timeout_s);
        }
        if (info) {
            info->status = r->status;
            info->t_submit = r->tSubmit;
            info->t_start = r->tStart;
            info->t_done = r->tDone;
        }
        code = r->status;
        msg = r->err;
        s->reqs.erase(it);
    });
    if (rc != MGG_OK) return rc;
    if (code != MGG_OK) return fail(code, "%s", msg.c_str());
    return MGG_OK;
}

int mgg_server_get_stats(mgg_server *s, mgg_server_stats *st)
{
    return guarded([&] {
        need_ptr(s, "server");
        need_ptr(st, "stats");
        std::lock_guard<std::mutex> lk(s->mu);
        st->submitted = s->submitted;
        st->completed = s->completed;
        st->failed = s->failedReqs;
        st->batches = s->batches;
        st->batched_directions = s->batchDirs;
        st->merged_batches = s->mergedBatches;
        st->budget_bytes = s->arena.size();
        st->resident_bytes = s->residentBytes;
        st->in_use_bytes = s->arena.used();
        st->peak_bytes = s->arena.peak();
        st->queued = (int)s->queue.size();
        st->in_flight = (int)s->inflight.size();
        st->volumes = (int)s->volumes.size();
    });
}

int mgg_server_hold(mgg_server *s, int hold)
{
    return guarded([&] {
        need_ptr(s, "server");
        std::lock_guard<std::mutex> lk(s->mu);
        s->hold = hold != 0;
        s->cvSched.notify_all();
    });
}

double mgg_server_clock(mgg_server *s) { return s ? s->now() : -1.0; }

int mgg_host_alloc(size_t bytes, void **ptr)
{
    return guarded([&] {
        need_ptr(ptr, "ptr");
        *ptr = nullptr;
        if (bytes == 0) raise(MGG_ERR_INVALID_ARGUMENT, "zero-byte pinned allocation");
        HIP_TRY(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    });
}

int mgg_host_free(void *ptr)
{
    return guarded([&] {
        if (ptr) HIP_TRY(hipHostFree(ptr));
    });
}

}  // extern "C"
