/*
 * cubam_amd: compressed BAM bytes -> per-record columns on AMD Instinct MI355X (gfx950).
 *
 * C ABI of libcubam_amd.so. Every pointer argument named *_d is a device
 * pointer (the caller allocates and frees all of them, e.g. from torch
 * tensors); `stream` is a hipStream_t. Every entry point only enqueues work on
 * `stream` and returns at once, except where it says it synchronises.
 * Return values: CBM_OK (0), or a negative CBM_ERR_* with a message for
 * cbm_last_error(). Data errors are reported per block / per record in
 * device memory (status arrays and result slots), never as return codes.
 *
 * Input padding: every compressed or inflated byte buffer passed in must be
 * readable for CBM_PAD bytes past its end (the kernels prefetch ahead).
 *
 * The reference (LouisFaure/cubam) has no C ABI of its own; on this path it
 * binds nvCOMP's batched Deflate C ABI through ctypes
 * (src/cubam/bgzf.py:161-178, call at :595-611) and launches NVRTC kernels
 * through CuPy (src/cubam/_kernels.py:637-702). The entry points below replace
 * those calls one for one, as cited on each.
 */
#ifndef CUBAM_AMD_H
#define CUBAM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CBM_PAD 64

enum {
    CBM_OK = 0,
    CBM_ERR_ARG = -1,    /* bad argument (null pointer, negative size, small workspace) */
    CBM_ERR_LAUNCH = -2, /* a HIP launch or runtime call failed */
};

/* ------------------------------------------------------------------ BGZF block scan (kernel 1) */

/* Result slots written by cbm_bgzf_scan (int64 each, device memory). */
enum {
    CBM_SCAN_NBLOCKS = 0,   /* blocks on the chain from offset 0, empty blocks included */
    CBM_SCAN_USED = 1,      /* bytes of buf those blocks cover; the rest is the carry */
    CBM_SCAN_STATUS = 2,    /* CBM_SCAN_OK or an error below */
    CBM_SCAN_TOTAL_OUT = 3, /* sum of isize over the blocks */
    CBM_SCAN_NCAND = 4,     /* header candidates found (diagnostic) */
    CBM_SCAN_RESULT_LEN = 8
};
enum {
    CBM_SCAN_OK = 0,
    CBM_SCAN_TRUNCATED = 1, /* eof and the chain does not end at n, or no block at all */
    CBM_SCAN_NO_BLOCK = 2,  /* no block header at offset 0 though there are others */
    CBM_SCAN_CORRUPT = 3,   /* isize > 65536, or an empty block that is not the EOF block */
    CBM_SCAN_CAP = 4        /* more blocks than cap */
};

/* Bytes of workspace cbm_bgzf_scan needs for a buffer of n bytes. */
int64_t cbm_bgzf_scan_workspace(int64_t n);

/*
 * Find the BGZF blocks of buf_d[0:n) (which starts at a block): header
 * candidates (the fixed gzip/BGZF header fields, BSIZE+1 >= 26, block inside
 * the buffer), the chain from offset 0, then validation. Writes, per block on
 * the chain (at most cap): blk_off, bsize, isize, the trailer CRC32 (crc) and
 * the output offset out_off (exclusive prefix sum of isize; cap + 1 entries,
 * out_off[n_blocks] = total); and the result slots above. buf_d must be
 * 16-byte aligned. Replaces append_offsets("bgzf_candidates") + bgzf_sizes +
 * chain + the checks of src/cubam/bgzf.py:549-584 (kernels at
 * src/cubam/_kernels.py:69-89, 662-702).
 */
int cbm_bgzf_scan(const uint8_t* buf_d, int64_t n, int eof, int64_t cap, int64_t* blk_off_d, int32_t* bsize_d,
                  int32_t* isize_d, uint32_t* crc_d, int64_t* out_off_d, int64_t* result_d, void* ws_d,
                  int64_t ws_bytes, void* stream);

/* ------------------------------------------------------------------ inflate + CRC32 (kernel 2) */

/* Per-block status written by cbm_inflate (low byte; bits 8-15 of a failed block name the
   check that failed, for diagnostics). */
enum {
    CBM_BLOCK_OK = 0,
    CBM_BLOCK_DEFLATE = 1, /* invalid DEFLATE stream (bad block type, code, distance, overrun) */
    CBM_BLOCK_SIZE = 2,    /* it does not inflate to isize bytes */
    CBM_BLOCK_CRC = 3      /* CRC32 of the output differs from the trailer */
};

/* Inflate kernel variants (cbm_inflate_tune.variant): a workgroup of NT threads decodes each
   block, each thread SEG bits of a round; or one block per lane (CBM_INFLATE_LANES); or the
   two-phase variants. A null tune selects CBM_INFLATE_TOK64_S512_SPLIT for CBM_INFLATE_SPLIT_MIN
   blocks or more, else CBM_INFLATE_TOK64_S512 (CBM_INFLATE_DEFAULT). */
enum {
    CBM_INFLATE_WG512_S128 = 0, /* NT 512, SEG 128: the one-phase kernel the others fall back to */
    CBM_INFLATE_LANES = 1,      /* blocks_per_wave lanes of a wavefront each inflate a block */
    CBM_INFLATE_WG128 = 2,      /* NT 128, SEG 256 */
    CBM_INFLATE_WG512 = 3,      /* NT 512, SEG 64 */
    CBM_INFLATE_WG64 = 4,       /* NT 64, SEG 512 */
    CBM_INFLATE_WG256_S128 = 5, /* NT 256, SEG 128 */
    CBM_INFLATE_WG128_S128 = 6, /* NT 128, SEG 128 */
    CBM_INFLATE_WG64_S128 = 7,  /* NT 64, SEG 128 */
    CBM_INFLATE_WG64_S256 = 8,  /* NT 64, SEG 256 */
    CBM_INFLATE_WG384_S128 = 9, /* NT 384, SEG 128 */
    CBM_INFLATE_WG256 = 10,     /* NT 256, SEG 256 */
    /* two-phase: a wavefront decodes each block in SEG-bit segments, literals straight to dst_d and
       matches to a list; a resolve pass applies the matches in dst_d; then the CRC32 check. Blocks
       with more matches than the first pass's lists hold go through both phases again with lists
       of the most matches a block can have; blocks either pass cannot finish (any failed check)
       go through CBM_INFLATE_WG512_S128 (same results) */
    CBM_INFLATE_TOK64_S512 = 11,
    CBM_INFLATE_TOK64_S1024 = 12,
    CBM_INFLATE_TOK64_S256 = 13,
    /* the two-phase inflate at SEG 512 with each block resolved by the wavefront that decoded it,
       right after (no resolve launch; blocks_per_wave and waves_per_group are not used) */
    CBM_INFLATE_TOK64_S512_FUSED = 14,
    /* CBM_INFLATE_TOK64_S512 with the blocks cut in two at 3/5: the first part decodes and resolves
       on `stream`, the second at the same time on a side stream of the library's (one per device
       and calling stream, default priority, kept for the life of the process), joined back into
       `stream` before the blocks the first phase left are decoded; a part's resolve overlaps the
       other's decode and the tails of the persistent launches. Under graph capture the fork and
       join are captured too, onto the side stream an earlier call on the same stream made; with
       none, it runs as CBM_INFLATE_TOK64_S512 on `stream` alone. */
    CBM_INFLATE_TOK64_S512_SPLIT = 15,
    CBM_INFLATE_VARIANTS = 16,
    CBM_INFLATE_DEFAULT = CBM_INFLATE_TOK64_S512 /* a null tune below CBM_INFLATE_SPLIT_MIN blocks */
};
/* blocks from which a null tune splits the two-phase inflate over two streams */
#define CBM_INFLATE_SPLIT_MIN 4096

/* Launch shape of the inflate kernel; zero fields select the variant's defaults. */
typedef struct {
    int32_t variant; /* CBM_INFLATE_* */
    /* CBM_INFLATE_LANES: lanes of a wavefront that each inflate a block, 1..64 (default 4);
       two-phase variants, the resolve pass's LDS layout: 1 or 2 = matches a lane per window of
       2 or 4 KiB (default 1); 4, 8, 16 or 32 = a history ring of that many KiB a wavefront, which
       takes each 2 KiB of output through LDS once and fuses the CRC32 into its write-back */
    int32_t blocks_per_wave;
    /* CBM_INFLATE_LANES: wavefronts per workgroup, 1..16 (default 2); two-phase variants: blocks
       (one a wavefront) per workgroup of the resolve pass: 2, 4 or 8 with 1 match a lane (default
       4, and 2 for CBM_INFLATE_TOK64_S512_SPLIT), 2 or 4 with 2; with a ring, 1, 2 or 4 (4 KiB,
       8 KiB), 1 or 2 (16 KiB), 1 (32 KiB) */
    int32_t waves_per_group;
    int32_t groups; /* workgroups of each decode launch; 0 = as many as the CUs hold */
} cbm_inflate_tune;

/* Bytes of workspace cbm_inflate needs for n_blocks blocks. */
int64_t cbm_inflate_workspace(int64_t n_blocks);

/*
 * Inflate n_blocks raw-DEFLATE BGZF payloads (src_d + blk_off[i] + 18, bsize[i]
 * - 26 bytes) into dst_d + out_off[i] (isize[i] bytes each), then check each
 * output's CRC32 against crc_expect[i] with one wavefront per block. Writes
 * crc_d[i] (CRC32 of what was written) and status_d[i]; result_d[0] receives
 * the number of blocks whose status is not CBM_BLOCK_OK.
 * Replaces nvcompBatchedDeflateDecompressAsync (src/cubam/bgzf.py:595-611)
 * and adds the CRC32 check the reference skips (docs/reading.md:48-51), using
 * the combination method of crc32_spans (src/cubam/_kernels.py:93-141).
 */
int cbm_inflate(const uint8_t* src_d, const int64_t* blk_off_d, const int32_t* bsize_d, const int32_t* isize_d,
                const uint32_t* crc_expect_d, const int64_t* out_off_d, int64_t n_blocks, uint8_t* dst_d,
                uint32_t* crc_d, int32_t* status_d, int64_t* result_d, const cbm_inflate_tune* tune, void* ws_d,
                int64_t ws_bytes, void* stream);

/*
 * cbm_inflate without a workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). No counterpart in the reference. It calls no
 * allocating routine: the block counter, statistics and match lists live in a
 * static device array of the library, 89.7 MB on each GPU it runs on, held
 * until the process ends, found once per device with hipGetSymbolAddress (the
 * calling thread's current device must be the stream's). It holds 512 match
 * lists, so tune->groups above 512 is rejected (CBM_ERR_ARG); the default
 * launch uses min(512, two a CU). A null or two-phase tune runs the one-phase
 * kernel (CBM_INFLATE_WG512_S128), whose workspace fits the array; results
 * (bytes, CRC32s, statuses) are those of cbm_inflate, every variant giving the
 * same. Every call on a GPU shares that GPU's array, and each call clears
 * what must start at zero: calls that overlap on one GPU (two streams, two
 * threads) race and must be serialised by the caller.
 */
int cbm_inflate_noalloc(const uint8_t* src_d, const int64_t* blk_off_d, const int32_t* bsize_d,
                        const int32_t* isize_d, const uint32_t* crc_expect_d, const int64_t* out_off_d,
                        int64_t n_blocks, uint8_t* dst_d, uint32_t* crc_d, int32_t* status_d, int64_t* result_d,
                        const cbm_inflate_tune* tune, void* stream);

/* CRC32 (zlib) of n spans data_d[off[i] : off[i] + len[i]], one wavefront per span. Replaces
   crc32_spans (src/cubam/_kernels.py:106-141). */
int cbm_crc32_spans(const uint8_t* data_d, const int64_t* off_d, const int64_t* len_d, int64_t n, uint32_t* out_d,
                    void* stream);

/* ------------------------------------------------------------------ record chain (kernel 3) */

/* Result slots written by cbm_record_chain (int64 each, device memory). */
enum {
    CBM_REC_N = 0,         /* complete records */
    CBM_REC_TAIL = 1,      /* offset (from data start) of the incomplete last record: the carry */
    CBM_REC_INVALID = 2,   /* records that fail the structural test (is_record) */
    CBM_REC_CIGAR = 3,     /* sum of n_cigar_op */
    CBM_REC_SEQ = 4,       /* sum of ceil(l_seq / 2) */
    CBM_REC_QUAL = 5,      /* sum of l_seq */
    CBM_REC_BAD_LINKS = 6, /* segments whose speculative walks did not chain (host repairs them) */
    CBM_REC_NSEG = 7,      /* segments */
    CBM_REC_RESULT_LEN = 8
};

/* Bytes of workspace for n inflated bytes; *seg_bytes receives the segment size used. */
int64_t cbm_record_chain_workspace(int64_t n, int64_t* seg_bytes);

/*
 * Find the record chain of data_d[start : n) (start is a record boundary) by
 * walking block_size from speculative per-segment starts and checking that the
 * walks link up. Writes the result slots and keeps per-segment state in ws_d
 * for cbm_record_offsets. Replaces append_offsets("record_candidates") +
 * record_ends + chain of src/cubam/bam.py:403-418 (kernels at
 * src/cubam/_kernels.py:250-289, 676-702).
 */
int cbm_record_chain(const uint8_t* data_d, int64_t start, int64_t n, int32_t n_ref, int64_t* result_d, void* ws_d,
                     int64_t ws_bytes, void* stream);

/*
 * Repair segments whose walks did not link (result slot CBM_REC_BAD_LINKS > 0):
 * walks the chain serially from start and rewrites the per-segment state.
 * Synchronises the stream. Rare: only data where the structural test finds a
 * false record start, or corrupt data.
 */
int cbm_record_chain_repair(const uint8_t* data_d, int64_t start, int64_t n, int32_t n_ref, int64_t* result_d,
                            void* ws_d, int64_t ws_bytes, void* stream);

/*
 * count_d[0] = the number of byte offsets o in [start, n) where a complete,
 * structurally valid record starts (is_record, one thread an offset). The
 * reader calls it only when the chain from start holds no record: the
 * reference then raises "No valid block or record..." when any candidate
 * exists and otherwise carries the bytes to the next chunk. Replaces the
 * count of append_offsets("record_candidates") (src/cubam/_kernels.py:280-284,
 * 662-673) as src/cubam/bam.py:407-412 and _kernels.py:687-689 use it.
 */
int cbm_record_candidates(const uint8_t* data_d, int64_t start, int64_t n, int32_t n_ref, int64_t* count_d,
                          void* stream);

/* Offsets (from data start) of the n_rec records found by cbm_record_chain, in order. */
int cbm_record_offsets(const uint8_t* data_d, int64_t start, int64_t n, const void* ws_d, int64_t ws_bytes,
                       int64_t* offsets_d, int64_t n_rec, void* stream);

/*
 * cbm_record_offsets, and with the offsets the CSR offset columns of the records' CIGAR words,
 * packed bases and qualities ([n_rec + 1] each, as cbm_scatter writes them; a NULL column is
 * skipped), from the sums the chain's walks kept: no pass over the records. For
 * cbm_scatter_csr. The reference computes them from the decoded batch (cumsum of n_cigar_op
 * and l_seq, src/cubam/bam.py:275-288).
 */
int cbm_record_offsets_csr(const uint8_t* data_d, int64_t start, int64_t n, const void* ws_d, int64_t ws_bytes,
                           int64_t* offsets_d, int64_t* cigar_off_d, int64_t* seq_off_d, int64_t* qual_off_d,
                           int64_t n_rec, void* stream);

/*
 * Record chains of many spans of one buffer in one launch sequence: the records of
 * span s are data_d[span_start_d[s] : span_end_d[s]), both ends record boundaries
 * (a batch of region requests, each inflated from a block where a record starts
 * and cut where the next request's first record would start). seg_base_d[s] is
 * the first segment of span s (seg_base_d[s + 1] - seg_base_d[s] =
 * ceil(span length / seg_bytes); n_seg = seg_base_d[n_spans]); span_nref_d[s] is
 * the reference count of the file span s comes from. Writes the result
 * slots (CBM_REC_TAIL unused; a span that does not end exactly at its end counts
 * in CBM_REC_INVALID) and span_rec_d[s] = index of the first record of span s,
 * span_rec_d[n_spans] = all records. No counterpart in the reference, which
 * decodes whole files (src/cubam/bam.py:421-446 is the single-span case).
 */
int64_t cbm_record_spans_workspace(int64_t n_seg);
int cbm_record_chain_spans(const uint8_t* data_d, const int64_t* span_start_d, const int64_t* span_end_d,
                           const int64_t* seg_base_d, const int32_t* span_nref_d, int64_t n_spans, int64_t n_seg,
                           int64_t* result_d, int64_t* span_rec_d, void* ws_d, int64_t ws_bytes, void* stream);
/* Repair of cbm_record_chain_spans as cbm_record_chain_repair (synchronises the stream). */
int cbm_record_chain_spans_repair(const uint8_t* data_d, const int64_t* span_start_d, const int64_t* span_end_d,
                                  const int64_t* seg_base_d, const int32_t* span_nref_d, int64_t n_spans,
                                  int64_t n_seg, int64_t* result_d, int64_t* span_rec_d, void* ws_d, int64_t ws_bytes,
                                  void* stream);
/* Offsets (from data start) of the n_rec records of all spans, span by span. */
int cbm_record_offsets_spans(const uint8_t* data_d, int64_t n_seg, const void* ws_d, int64_t ws_bytes,
                             int64_t* offsets_d, int64_t n_rec, void* stream);
/* cbm_record_offsets_spans with the CSR offset columns, as cbm_record_offsets_csr. */
int cbm_record_offsets_spans_csr(const uint8_t* data_d, int64_t n_seg, const void* ws_d, int64_t ws_bytes,
                                 int64_t* offsets_d, int64_t* cigar_off_d, int64_t* seq_off_d, int64_t* qual_off_d,
                                 int64_t n_rec, void* stream);

/* ------------------------------------------------------------------ column scatter (kernel 4) */

/* Column outputs of cbm_scatter (device pointers; a NULL pointer skips that column). */
typedef struct {
    int32_t *ref, *pos, *end;
    uint8_t* mapq;
    uint16_t* flag;
    int32_t *nref, *npos, *tlen, *l_seq;
    uint16_t *clip_left, *clip_right;
    int64_t* cigar_off; /* [n + 1] */
    uint32_t* cigar;    /* raw CIGAR words (len << 4 | op) */
    int64_t* seq_off;   /* [n + 1] */
    uint8_t* seq;       /* 4-bit packed bases, BAM order (high nibble first) */
    int64_t* qual_off;  /* [n + 1] */
    uint8_t* qual;
    /* aux tags: up to 4 names, each type u8 / value offset i64 / length i32 / value i64 per record */
    int32_t n_tags;
    char tag_names[8];
    uint8_t* tag_type[4];
    int64_t* tag_offset[4];
    int32_t* tag_length[4];
    int64_t* tag_value[4];
} cbm_columns;

/* Bytes of workspace cbm_scatter needs for n_rec records. */
int64_t cbm_scatter_workspace(int64_t n_rec);

/*
 * Decode the records at offsets_d[0:n_rec) of data_d into columns: the 11
 * fixed fields of decode_fixed (src/cubam/_kernels.py:292-319), the CSR
 * CIGAR / seq / qual columns, and aux tags as find_tag
 * (src/cubam/_kernels.py:334-388). Replaces RecordBatch.decode
 * (src/cubam/bam.py:268-288) and Tag.__init__ (src/cubam/bam.py:184-208).
 */
int cbm_scatter(const uint8_t* data_d, const int64_t* offsets_d, int64_t n_rec, const cbm_columns* cols, void* ws_d,
                int64_t ws_bytes, void* stream);

/*
 * cbm_scatter for records whose CSR offset columns are already written (cols->cigar_off,
 * seq_off and qual_off are inputs here, from cbm_record_offsets_csr for the same records):
 * needs no workspace. Replaces the same RecordBatch.decode (src/cubam/bam.py:268-288).
 */
int cbm_scatter_csr(const uint8_t* data_d, const int64_t* offsets_d, int64_t n_rec, const cbm_columns* cols,
                    void* stream);

/* ------------------------------------------------------------------ on-demand record kernels */
/* One thread per record, reading records in place at data_d + off_d[i] (batch offsets). */

/* FNV-1a 64 (basis 1469598103934665603, as the reference) of each read name. Replaces the
   name_hash kernel of RecordBatch.name_hash (src/cubam/bam.py:322-330, _kernels.py:322-330). */
int cbm_name_hash(const uint8_t* data_d, const int64_t* off_d, int64_t n, uint64_t* out_d, void* stream);

/* FNV-1a 64 of spans data_d[voff[i] : + vlen[i]), 0 where vlen[i] < 0. Replaces hash_spans of
   Tag.hash (src/cubam/bam.py:221-230, _kernels.py:533-541). */
int cbm_hash_spans(const uint8_t* data_d, const int64_t* voff_d, const int32_t* vlen_d, int64_t n, uint64_t* out_d,
                   void* stream);

/* Values [ACGT]{1,24}(-[0-9]{1,3})? packed into uint64 codes (2 bits a base, base count at bit
   48, suffix + 1 at bit 53); ok_d[i] = 0 and code 0 where a value does not fit. Replaces
   pack_acgt of Tag.packed (src/cubam/bam.py:232-248, _kernels.py:543-571). */
int cbm_pack_acgt(const uint8_t* data_d, const int64_t* voff_d, const int32_t* vlen_d, int64_t n, uint64_t* out_d,
                  uint8_t* ok_d, void* stream);

/* Aligned bases (CIGAR M, =, X) of each read. Replaces aligned_len of
   RecordBatch.aligned_length (src/cubam/bam.py:341-348, _kernels.py:390-403). */
int cbm_aligned_length(const uint8_t* data_d, const int64_t* off_d, int64_t n, int32_t* out_d, void* stream);

/* Aligned blocks of each read (runs of M/=/X, and D unless split_deletions, between N skips):
   with fill = 0, cnt_d[i] = blocks of read i; with fill = 1, blocks written from first_d[i] as
   (rec, start, end). Replaces blocks of RecordBatch.blocks (src/cubam/bam.py:350-373,
   _kernels.py:405-437). */
int cbm_blocks(const uint8_t* data_d, const int64_t* off_d, int64_t n, int split_deletions, int fill,
               const int64_t* first_d, int32_t* cnt_d, int64_t* rec_d, int32_t* start_d, int32_t* end_d,
               void* stream);

/* Base of read rec_d[j] at reference position refpos_d[j]: 0..3 (A C G T), 4 other, -1 deletion,
   -2 reference skip, -3 not covered. Replaces base_at of RecordBatch.base_at
   (src/cubam/bam.py:375-389, _kernels.py:498-529). */
int cbm_base_at(const uint8_t* data_d, const int64_t* off_d, const int64_t* rec_d, const int32_t* refpos_d, int64_t m,
                int8_t* out_d, void* stream);

/* ------------------------------------------------------------------ BGZF compression (kernel 5) */

/* Output bytes reserved per block by cbm_deflate (a BGZF block is at most 65,536 bytes). */
#define CBM_DEFLATE_SLOT (65536 + 256)
/* Payload bytes a block may hold (htslib's BGZF_BLOCK_SIZE cut). */
#define CBM_DEFLATE_MAX_IN 65280

/* Bytes of workspace cbm_deflate needs for n_blocks blocks. */
int64_t cbm_deflate_workspace(int64_t n_blocks);

/*
 * Compress n_blocks payloads src_d[in_off[i] : + in_len[i]) (each at most
 * CBM_DEFLATE_MAX_IN bytes) into whole BGZF blocks (gzip header with the BC
 * extra field, raw DEFLATE, CRC32 and ISIZE trailer), block i at
 * slots_d + i * CBM_DEFLATE_SLOT; bsize_d[i] receives its size in bytes and
 * crc_d[i] (may be NULL) the payload's CRC32. Greedy level-1 matching and
 * dynamic Huffman codes; a block that does not shrink is stored. Replaces
 * _bgzf_compress (src/cubam/bgzf.py:380-422: nvCOMP Deflate, crc32_spans and
 * the bgzf_frame kernel of src/cubam/_kernels.py:161-191).
 */
int cbm_deflate(const uint8_t* src_d, const int64_t* in_off_d, const int32_t* in_len_d, int64_t n_blocks,
                uint8_t* slots_d, int32_t* bsize_d, uint32_t* crc_d, void* ws_d, int64_t ws_bytes, void* stream);

/* Copy n spans src_d[src_off[i] : + len[i]) to dst_d + dst_off[i], one wavefront per span
   (record gathers and block packing). Replaces copy_spans (src/cubam/_kernels.py:34-43, 644-659). */
int cbm_copy_spans(const uint8_t* src_d, const int64_t* src_off_d, const int64_t* len_d, const int64_t* dst_off_d,
                   int64_t n, uint8_t* dst_d, void* stream);

/* ------------------------------------------------------------------ misc */

/* Copy the last error message of this thread into buf (NUL-terminated); returns its length. */
int cbm_last_error(char* buf, int64_t len);

/* Library version string. */
const char* cbm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CUBAM_AMD_H */
