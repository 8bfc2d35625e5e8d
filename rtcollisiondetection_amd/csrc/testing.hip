// Test entry points: two building blocks of the discrete check run alone on caller buffers, so that
// they can be checked the way the reference checks its own (RTCD/Utils/Test/src/testExclusiveScan.cpp,
// testStreamCompact.cpp): the wave64 exclusive scan against a CPU exclusive prefix sum, and the
// candidate compaction against a CPU filter.  Not used by any check; they allocate what they need.
#include "ctx.hpp"

extern "C" {

rtcd_status rtcd_test_exclusive_scan(const uint32_t* d_in, uint32_t* d_out, uint32_t n, uint32_t* d_total) {
    if ((n && (!d_in || !d_out)) || !d_total) return fail(RTCD_ERR_INVALID_ARG, "null buffer");
    return guarded([&]() -> rtcd_status {
        hipStream_t s = nullptr;
        DevBuf<uint32_t> tmp(scan_tmp_words(n));
        exclusive_scan(d_in, d_out, n, d_total, tmp.p, s);
        RTCD_HIP(hipStreamSynchronize(s));
        return RTCD_OK;
    });
}

rtcd_status rtcd_test_compact(const uint32_t* d_records, uint32_t n, const uint32_t* d_flags, uint32_t n_links,
                              int32_t keep_flagged, uint32_t* d_out, uint32_t* d_count) {
    if ((n && (!d_records || !d_flags || !d_out)) || !d_count || n_links == 0)
        return fail(RTCD_ERR_INVALID_ARG, "null buffer or zero links");
    return guarded([&]() -> rtcd_status {
        hipStream_t s = nullptr;
        // the kernel reads the list length from the gather prefix's last word
        std::vector<uint32_t> pre(kCandSegments + 1, 0u);
        pre[kCandSegments] = n;
        DevBuf<uint32_t> d_pre;
        d_pre.upload(pre, s);
        fill_u32(d_count, 1, 0u, s);
        DiscreteArgs a{};
        a.n_links = (int) n_links;
        a.flags = const_cast<uint32_t*>(d_flags);
        a.flat = reinterpret_cast<uint2*>(const_cast<uint32_t*>(d_records));
        a.seg_pre = d_pre.p;
        a.r2o_list = reinterpret_cast<uint2*>(d_out);
        a.r2o_count = d_count;
        launch_compact_unflagged(a, s, keep_flagged != 0);
        RTCD_HIP(hipStreamSynchronize(s));
        return RTCD_OK;
    });
}

}  // extern "C"
