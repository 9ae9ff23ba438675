/*
 * The oracle: a CPU restatement of the reference's read path, for tests and
 * the CPU baseline only. It is never part of the product path.
 *
 * Reference files followed (paths relative to the reference repository):
 *   BGZF block chain + validation  src/cubam/bgzf.py:542-584 and the header
 *                                  predicate src/cubam/_kernels.py:69-89
 *   Inflate                        src/cubam/bgzf.py:587-611 (nvCOMP there;
 *                                  here zlib 1.2.11 raw inflate), plus the CRC32
 *                                  check the reference does not do
 *                                  (docs/reading.md:48-51)
 *   Record chain                   src/cubam/bam.py:403-418 and is_record
 *                                  src/cubam/_kernels.py:235-289
 *   Fixed columns                  src/cubam/_kernels.py:292-319 (decode_fixed)
 *   Tags                           src/cubam/_kernels.py:334-388 (find_tag)
 *   CIGAR / seq / qual columns     raw slices of each record (SAMv1 4.2)
 *
 * All entry points are plain C with caller-allocated buffers.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

static inline uint32_t u32(const uint8_t* b, int64_t o) {
    return (uint32_t)b[o] | ((uint32_t)b[o + 1] << 8) | ((uint32_t)b[o + 2] << 16) | ((uint32_t)b[o + 3] << 24);
}
static inline uint32_t u16(const uint8_t* b, int64_t o) { return (uint32_t)b[o] | ((uint32_t)b[o + 1] << 8); }

/* ------------------------------------------------------------------ BGZF */

/* _kernels.py:72-79: the fixed header fields, BSIZE + 1 >= 26 and the block inside the buffer */
int orc_bgzf_candidate(const uint8_t* b, int64_t n, int64_t o) {
    if (o + 18 > n) return 0;
    if (!(b[o] == 31 && b[o + 1] == 139 && b[o + 2] == 8 && b[o + 3] == 4 && b[o + 10] == 6 && b[o + 11] == 0 &&
          b[o + 12] == 66 && b[o + 13] == 67 && b[o + 14] == 2 && b[o + 15] == 0))
        return 0;
    int64_t bs = (int64_t)u16(b, o + 16) + 1;
    return bs >= 26 && o + bs <= n;
}

enum { ORC_OK = 0, ORC_TRUNCATED = 1, ORC_NO_BLOCK = 2, ORC_CORRUPT = 3, ORC_CAP = 4 };

/*
 * Blocks of buf (starting at a block), as bgzf.py:542-584 keeps them:
 * candidates at every offset, the chain from offset 0 through candidates, then
 * "used" = end of the chain. With eof the chain must end at n. Empty blocks
 * must be the 28-byte empty block (payload 03 00); isize <= 65536. Empty blocks
 * are dropped from the output.
 * out[0] = blocks kept, out[1] = used, out[2] = status (ORC_*), out[3] = blocks in the chain.
 */
void orc_bgzf_scan(const uint8_t* b, int64_t n, int eof, int64_t* blk_off, int32_t* bsize, int32_t* isize,
                   uint32_t* crc, int64_t cap, int64_t* out) {
    out[0] = out[1] = out[3] = 0;
    out[2] = ORC_OK;
    int64_t first = -1;
    for (int64_t o = 0; o < n; o++)
        if (orc_bgzf_candidate(b, n, o)) { first = o; break; }
    if (first < 0) {
        if (eof && n) out[2] = ORC_TRUNCATED;
        return;
    }
    if (first != 0) { out[2] = ORC_NO_BLOCK; return; }
    int64_t o = 0, kept = 0, chained = 0;
    int bad = 0;
    while (o < n && orc_bgzf_candidate(b, n, o)) {
        int64_t bs = (int64_t)u16(b, o + 16) + 1;
        uint32_t is = u32(b, o + bs - 4);
        chained++;
        if (is == 0) {
            if (bs != 28 || b[o + 18] != 3 || b[o + 19] != 0) bad = 1;
        } else if (is > 65536) {
            bad = 1;
        } else {
            if (kept >= cap) { out[2] = ORC_CAP; return; }
            blk_off[kept] = o;
            bsize[kept] = (int32_t)bs;
            isize[kept] = (int32_t)is;
            crc[kept] = u32(b, o + bs - 8);
            kept++;
        }
        o += bs;
    }
    out[1] = o;
    out[3] = chained;
    if (eof && o != n) { out[2] = ORC_TRUNCATED; return; }
    if (bad) { out[2] = ORC_CORRUPT; return; }
    out[0] = kept;
}

/* ------------------------------------------------------------------ inflate */
enum { ORC_INFLATE_OK = 0, ORC_INFLATE_ERROR = 1, ORC_INFLATE_SIZE = 2, ORC_INFLATE_CRC = 3 };

/* Raw DEFLATE payload of one block (bytes 18 .. bsize-8) into dst (isize bytes); status per enum. */
static int inflate_one(const uint8_t* blk, int32_t bs, uint8_t* dst, int32_t is, uint32_t* crc_out) {
    z_stream z;
    memset(&z, 0, sizeof z);
    if (inflateInit2(&z, -15) != Z_OK) return ORC_INFLATE_ERROR;
    z.next_in = (Bytef*)(blk + 18);
    z.avail_in = (uInt)(bs - 26);
    z.next_out = dst;
    z.avail_out = (uInt)is;
    int rc = inflate(&z, Z_FINISH);
    int64_t got = (int64_t)z.total_out;
    inflateEnd(&z);
    uint32_t c = (uint32_t)crc32(0L, dst, (uInt)got);
    *crc_out = c;
    if (rc != Z_STREAM_END) {
        /* output full before the end of the stream: more data than the trailer says */
        return (rc == Z_BUF_ERROR && got == is) ? ORC_INFLATE_SIZE : ORC_INFLATE_ERROR;
    }
    if (got != is) return ORC_INFLATE_SIZE;
    if (c != u32(blk, bs - 8)) return ORC_INFLATE_CRC;
    return ORC_INFLATE_OK;
}

typedef struct {
    const uint8_t* src;
    const int64_t* blk_off;
    const int32_t* bsize;
    const int32_t* isize;
    const int64_t* out_off;
    uint8_t* dst;
    uint32_t* crc;
    int32_t* status;
    int64_t n, next;
    pthread_mutex_t mu;
} inf_job;

static void* inf_worker(void* a) {
    inf_job* j = (inf_job*)a;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        int64_t i = j->next;
        j->next += 16;
        pthread_mutex_unlock(&j->mu);
        if (i >= j->n) return NULL;
        int64_t e = i + 16 < j->n ? i + 16 : j->n;
        for (; i < e; i++)
            j->status[i] = inflate_one(j->src + j->blk_off[i], j->bsize[i], j->dst + j->out_off[i], j->isize[i],
                                       &j->crc[i]);
    }
}

/* Inflate n blocks of src into dst at out_off[i]; crc[i] = CRC32 of the output, status[i] per enum. */
void orc_inflate(const uint8_t* src, const int64_t* blk_off, const int32_t* bsize, const int32_t* isize,
                 const int64_t* out_off, int64_t n, uint8_t* dst, uint32_t* crc, int32_t* status, int n_threads) {
    inf_job j = {src, blk_off, bsize, isize, out_off, dst, crc, status, n, 0, PTHREAD_MUTEX_INITIALIZER};
    if (n_threads <= 1) {
        inf_worker(&j);
        return;
    }
    if (n_threads > 256) n_threads = 256;
    pthread_t th[256];
    for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, inf_worker, &j);
    for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
}

/* zlib's CRC32 of a span (for tests) */
uint32_t orc_crc32(const uint8_t* p, int64_t n) { return (uint32_t)crc32(0L, p, (uInt)n); }

/* ------------------------------------------------------------------ records */

/* SAMv1 5.3 (and _kernels.py:235-243) */
int orc_reg2bin(int beg, int end) {
    --end;
    if (beg >> 14 == end >> 14) return ((1 << 15) - 1) / 7 + (beg >> 14);
    if (beg >> 17 == end >> 17) return ((1 << 12) - 1) / 7 + (beg >> 17);
    if (beg >> 20 == end >> 20) return ((1 << 9) - 1) / 7 + (beg >> 20);
    if (beg >> 23 == end >> 23) return ((1 << 6) - 1) / 7 + (beg >> 23);
    if (beg >> 26 == end >> 26) return ((1 << 3) - 1) / 7 + (beg >> 26);
    return 0;
}

static inline int ref_consuming(uint32_t op) { return op == 0 || op == 2 || op == 3 || op == 7 || op == 8; }

/* _kernels.py:250-278: a complete, structurally valid record starts at o */
int orc_is_record(const uint8_t* b, int64_t n, int64_t o, int n_ref) {
    if (o + 36 > n) return 0;
    int32_t bs = (int32_t)u32(b, o);
    if (bs < 33 || bs > (1 << 22) || o + 4 + bs > n) return 0;
    int32_t ref = (int32_t)u32(b, o + 4), pos = (int32_t)u32(b, o + 8);
    int32_t nref = (int32_t)u32(b, o + 24), npos = (int32_t)u32(b, o + 28);
    if (ref < -1 || ref >= n_ref || pos < -1 || nref < -1 || nref >= n_ref || npos < -1) return 0;
    uint32_t lrn = b[o + 12], ncig = u16(b, o + 16), flag = u16(b, o + 18), lseq = u32(b, o + 20);
    if (lrn < 2 || flag >= 4096 || lseq > (1u << 22)) return 0;
    if (32LL + lrn + 4LL * ncig + (lseq + 1) / 2 + lseq > bs) return 0;
    int64_t nm = o + 36;
    if (b[nm + lrn - 1] != 0) return 0;
    for (uint32_t i = 0; i + 1 < lrn; i++) {
        uint8_t c = b[nm + i];
        if (c < 33 || c > 126 || c == '@') return 0;
    }
    int64_t cg = nm + lrn, rlen = 0;
    for (uint32_t i = 0; i < ncig; i++) {
        uint32_t v = u32(b, cg + 4LL * i), op = v & 15;
        if (op > 8) return 0;
        if (ref_consuming(op)) rlen += v >> 4;
    }
    if (pos >= 0) {
        int end = pos + (int)(rlen > 0 ? rlen : 1);
        if (u16(b, o + 14) != (uint32_t)orc_reg2bin(pos, end)) return 0;
    }
    return 1;
}

/*
 * The record chain of data (starting at a record), as bam.py:403-418 and
 * chain (_kernels.py:676-702) keep it: from offset 0, each record that fits
 * and passes is_record, then the next at its end. The chain stops at the
 * first record that does not fit (the incomplete last one) or fails the test;
 * out[1] (the tail, carried to the next chunk) is where that record starts.
 * out[0] = records, out[2] = status: 0 ok, 2 = the chain stopped at a complete
 * record that fails is_record, 4 = cap exceeded.
 */
void orc_record_chain(const uint8_t* b, int64_t n, int n_ref, int64_t* off, int64_t cap, int64_t* out) {
    int64_t o = 0, k = 0;
    out[2] = 0;
    while (o + 4 <= n) {
        int64_t bs = (int64_t)u32(b, o);
        if (o + 4 + bs > n) break;
        if (!orc_is_record(b, n, o, n_ref)) { out[2] = 2; break; }
        if (k >= cap) { out[2] = 4; break; }
        off[k++] = o;
        o += 4 + bs;
    }
    out[0] = k;
    out[1] = o;
}

/*
 * Offsets of b[0:n) where orc_is_record holds, counted by brute force: the
 * count append_offsets("record_candidates") makes (_kernels.py:280-284,
 * 662-673). bam.py:407-412 and chain (_kernels.py:687-689) use it when no
 * record starts at offset 0: with any candidate the reference raises "No
 * valid block or record...", with none it carries the data to the next chunk.
 */
int64_t orc_record_candidates(const uint8_t* b, int64_t n, int n_ref) {
    int64_t c = 0;
    for (int64_t o = 0; o < n; o++) c += orc_is_record(b, n, o, n_ref);
    return c;
}

/* Sizes that the CSR columns need: n_cigar, ceil(l_seq / 2), l_seq per record. */
void orc_var_sizes(const uint8_t* b, const int64_t* off, int64_t n, int32_t* ncig, int32_t* lseq) {
    for (int64_t i = 0; i < n; i++) {
        ncig[i] = (int32_t)u16(b, off[i] + 16);
        lseq[i] = (int32_t)u32(b, off[i] + 20);
    }
}

/*
 * decode_fixed (_kernels.py:292-319): 11 columns. end = pos + reference-consuming
 * length (no +1 for CIGAR-less reads); a soft clip at op 0 is the left clip, any
 * later S the right clip; clips saturate at 65535.
 * Then the CSR columns: cigar words, packed seq bytes, qual bytes, at the
 * offsets given (exclusive prefix sums of n_cigar, ceil(l_seq/2), l_seq).
 */
void orc_decode_columns(const uint8_t* b, const int64_t* off, int64_t n, int32_t* ref, int32_t* pos, int32_t* end,
                        uint8_t* mapq, uint16_t* flag, int32_t* nref, int32_t* npos, int32_t* tlen, int32_t* l_seq,
                        uint16_t* clip_left, uint16_t* clip_right, const int64_t* cigar_off, uint32_t* cigar,
                        const int64_t* seq_off, uint8_t* seq, const int64_t* qual_off, uint8_t* qual) {
    for (int64_t i = 0; i < n; i++) {
        int64_t o = off[i];
        int32_t p = (int32_t)u32(b, o + 8);
        ref[i] = (int32_t)u32(b, o + 4);
        pos[i] = p;
        mapq[i] = b[o + 13];
        flag[i] = (uint16_t)u16(b, o + 18);
        nref[i] = (int32_t)u32(b, o + 24);
        npos[i] = (int32_t)u32(b, o + 28);
        tlen[i] = (int32_t)u32(b, o + 32);
        uint32_t ls = u32(b, o + 20);
        l_seq[i] = (int32_t)ls;
        uint32_t lrn = b[o + 12], ncig = u16(b, o + 16);
        int64_t cg = o + 36 + lrn, rlen = 0;
        uint32_t cl = 0, cr = 0;
        for (uint32_t k = 0; k < ncig; k++) {
            uint32_t v = u32(b, cg + 4LL * k), op = v & 15, len = v >> 4;
            if (ref_consuming(op)) rlen += len;
            if (op == 4) {
                if (k == 0) cl = len;
                else cr = len;
            }
            if (cigar) cigar[cigar_off[i] + k] = v;
        }
        end[i] = p + (int32_t)rlen;
        clip_left[i] = (uint16_t)(cl > 65535 ? 65535 : cl);
        clip_right[i] = (uint16_t)(cr > 65535 ? 65535 : cr);
        int64_t sq = cg + 4LL * ncig;
        if (seq) memcpy(seq + seq_off[i], b + sq, (ls + 1) / 2);
        if (qual) memcpy(qual + qual_off[i], b + sq + (ls + 1) / 2, ls);
    }
}

/*
 * find_tag (_kernels.py:334-388): type 0 = missing, 255 = malformed aux data,
 * else the type character; Z/H value offset and length without the NUL; B
 * length from the subtype byte to the end; integer value of c/C/s/S/i/I/A
 * (f gives 0).
 */
void orc_find_tag(const uint8_t* b, const int64_t* off, int64_t n, uint8_t t0, uint8_t t1, uint8_t* type,
                  int64_t* voff, int32_t* vlen, int64_t* ival) {
    for (int64_t i = 0; i < n; i++) {
        int64_t o = off[i];
        int64_t rec_end = o + 4 + (int64_t)u32(b, o);
        uint32_t lrn = b[o + 12], ncig = u16(b, o + 16), lseq = u32(b, o + 20);
        int64_t a = o + 36 + lrn + 4LL * ncig + (lseq + 1) / 2 + lseq;
        uint8_t ty_out = 0;
        int64_t vo = 0, iv = 0;
        int32_t vl = 0;
        while (a + 3 <= rec_end) {
            uint8_t ty = b[a + 2];
            int64_t v = a + 3, nx;
            switch (ty) {
                case 'A': case 'c': case 'C': nx = v + 1; break;
                case 's': case 'S': nx = v + 2; break;
                case 'i': case 'I': case 'f': nx = v + 4; break;
                case 'Z': case 'H':
                    nx = v;
                    while (nx < rec_end && b[nx] != 0) nx++;
                    nx++;
                    break;
                case 'B': {
                    if (v + 5 > rec_end) { nx = -1; break; }
                    uint8_t sub = b[v];
                    int64_t cnt = (int64_t)u32(b, v + 1);
                    int sz = (sub == 'c' || sub == 'C') ? 1 : (sub == 's' || sub == 'S') ? 2 : 4;
                    nx = v + 5 + cnt * sz;
                    break;
                }
                default: nx = -1;
            }
            if (nx < 0 || nx > rec_end) { ty_out = 255; break; }
            if (b[a] == t0 && b[a + 1] == t1) {
                ty_out = ty;
                vo = v;
                switch (ty) {
                    case 'c': iv = (int8_t)b[v]; break;
                    case 'C': case 'A': iv = b[v]; break;
                    case 's': iv = (int16_t)u16(b, v); break;
                    case 'S': iv = u16(b, v); break;
                    case 'i': iv = (int32_t)u32(b, v); break;
                    case 'I': iv = u32(b, v); break;
                    case 'Z': case 'H': vl = (int32_t)(nx - 1 - v); break;
                    case 'B': vl = (int32_t)(nx - v); break;
                    default: break;
                }
                break;
            }
            a = nx;
        }
        type[i] = ty_out;
        voff[i] = vo;
        vlen[i] = vl;
        ival[i] = iv;
    }
}

/* FNV-1a 64 (bam.py:164-173, _kernels.py:322-330, 532-540) */
uint64_t orc_fnv1a(const uint8_t* p, int64_t n) {
    uint64_t x = 1469598103934665603ULL;
    for (int64_t k = 0; k < n; k++) { x ^= p[k]; x *= 1099511628211ULL; }
    return x;
}

/* pack_acgt (_kernels.py:544-571): [ACGT]{1,24}(-[0-9]{1,3})? -> code, ok */
uint64_t orc_pack_acgt(const uint8_t* p, int32_t n, int* ok) {
    int64_t k = 0;
    uint64_t code = 0, suffix = 0;
    uint32_t nb = 0;
    int good = n > 0;
    for (; good && k < n && p[k] != '-'; k++) {
        uint8_t c = p[k];
        uint32_t x = c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : 4;
        if (x == 4 || nb == 24) good = 0;
        else code |= (uint64_t)x << (2 * nb++);
    }
    if (good && k < n) {
        k++;
        uint32_t nd = 0;
        for (; k < n; k++, nd++) {
            if (p[k] < '0' || p[k] > '9' || nd == 3) { good = 0; break; }
            suffix = suffix * 10 + (p[k] - '0');
        }
        if (nd == 0) good = 0;
        suffix += 1;
    }
    if (nb == 0) good = 0;
    *ok = good;
    return good ? (code | ((uint64_t)nb << 48) | (suffix << 53)) : 0;
}

/* ------------------------------------------------------------------ whole-chunk decode for the CPU baseline */
typedef struct {
    const uint8_t* data;
    const int64_t* off;
    int64_t n;
    int32_t *ref, *pos, *end, *nref, *npos, *tlen, *l_seq;
    uint8_t* mapq;
    uint16_t *flag, *cl, *cr;
    const int64_t *cig_off, *seq_off, *qual_off;
    uint32_t* cigar;
    uint8_t *seq, *qual;
    uint8_t* t_type[2];
    int64_t* t_off[2];
    int32_t* t_len[2];
    int64_t* t_val[2];
    const char* tags;
    int n_tags;
    int64_t next;
    pthread_mutex_t mu;
} col_job;

static void* col_worker(void* a) {
    col_job* j = (col_job*)a;
    const int64_t step = 8192;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        int64_t i = j->next;
        j->next += step;
        pthread_mutex_unlock(&j->mu);
        if (i >= j->n) return NULL;
        int64_t m = i + step < j->n ? step : j->n - i;
        orc_decode_columns(j->data, j->off + i, m, j->ref + i, j->pos + i, j->end + i, j->mapq + i, j->flag + i,
                           j->nref + i, j->npos + i, j->tlen + i, j->l_seq + i, j->cl + i, j->cr + i, j->cig_off + i,
                           j->cigar, j->seq_off + i, j->seq, j->qual_off + i, j->qual);
        for (int t = 0; t < j->n_tags; t++)
            orc_find_tag(j->data, j->off + i, m, (uint8_t)j->tags[2 * t], (uint8_t)j->tags[2 * t + 1],
                         j->t_type[t] + i, j->t_off[t] + i, j->t_len[t] + i, j->t_val[t] + i);
    }
}

/* Fill every column of n records with n_threads threads (CSR offsets given). tags: "CBUB" style, up to 2. */
void orc_columns_mt(const uint8_t* data, const int64_t* off, int64_t n, int32_t* ref, int32_t* pos, int32_t* end,
                    uint8_t* mapq, uint16_t* flag, int32_t* nref, int32_t* npos, int32_t* tlen, int32_t* l_seq,
                    uint16_t* cl, uint16_t* cr, const int64_t* cig_off, uint32_t* cigar, const int64_t* seq_off,
                    uint8_t* seq, const int64_t* qual_off, uint8_t* qual, const char* tags, int n_tags,
                    uint8_t* t_type, int64_t* t_off, int32_t* t_len, int64_t* t_val, int n_threads) {
    col_job j;
    memset(&j, 0, sizeof j);
    j.data = data; j.off = off; j.n = n;
    j.ref = ref; j.pos = pos; j.end = end; j.mapq = mapq; j.flag = flag; j.nref = nref; j.npos = npos;
    j.tlen = tlen; j.l_seq = l_seq; j.cl = cl; j.cr = cr;
    j.cig_off = cig_off; j.cigar = cigar; j.seq_off = seq_off; j.seq = seq; j.qual_off = qual_off; j.qual = qual;
    j.tags = tags;
    j.n_tags = n_tags > 2 ? 2 : n_tags;
    for (int t = 0; t < j.n_tags; t++) {
        j.t_type[t] = t_type + t * n;
        j.t_off[t] = t_off + t * n;
        j.t_len[t] = t_len + t * n;
        j.t_val[t] = t_val + t * n;
    }
    pthread_mutex_init(&j.mu, NULL);
    if (n_threads <= 1) {
        col_worker(&j);
        return;
    }
    if (n_threads > 256) n_threads = 256;
    pthread_t th[256];
    for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, col_worker, &j);
    for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
}
