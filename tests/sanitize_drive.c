/*
 * Host C of the test infrastructure under AddressSanitizer and UndefinedBehaviorSanitizer: the
 * synthetic BAM generator (cubam_amd/synth/gen.c) and the oracle (oracle/cpu_decode.c), driven
 * over valid files of both recipes and over corrupted, truncated and random bytes. Built and run
 * by tests/test_sanitize.py (gcc, -fsanitize=address,undefined); any report aborts the run.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* gen.c */
void* cbs_open(int kind, uint64_t seed, int64_t n_total);
void cbs_close(void* h);
int64_t cbs_header(void* h, uint8_t* out, int64_t cap);
int64_t cbs_records(void* h, uint8_t* out, int64_t cap, int64_t max_rec, int64_t* n_rec);
int cbs_bgzf_blocks_codec(const uint8_t* src, const int64_t* off, const int32_t* len, int64_t n,
                          const int8_t* level, const uint8_t* fixed, uint8_t* dst, int64_t stride, int32_t* bsize,
                          int n_threads, int codec);
int cbs_have_libdeflate(void);
/* cpu_decode.c */
void orc_bgzf_scan(const uint8_t* b, int64_t n, int eof, int64_t* blk_off, int32_t* bsize, int32_t* isize,
                   uint32_t* crc, int64_t cap, int64_t* out);
void orc_inflate(const uint8_t* src, const int64_t* blk_off, const int32_t* bsize, const int32_t* isize,
                 const int64_t* out_off, int64_t n, uint8_t* dst, uint32_t* crc, int32_t* status, int n_threads);
int orc_is_record(const uint8_t* b, int64_t n, int64_t o, int n_ref);
void orc_record_chain(const uint8_t* b, int64_t n, int n_ref, int64_t* off, int64_t cap, int64_t* out);
int64_t orc_record_candidates(const uint8_t* b, int64_t n, int n_ref);
void orc_var_sizes(const uint8_t* b, const int64_t* off, int64_t n, int32_t* ncig, int32_t* lseq);
void orc_find_tag(const uint8_t* b, const int64_t* off, int64_t n, uint8_t t0, uint8_t t1, uint8_t* type,
                  int64_t* voff, int32_t* vlen, int64_t* ival);
uint64_t orc_fnv1a(const uint8_t* p, int64_t n);
uint64_t orc_pack_acgt(const uint8_t* p, int32_t n, int* ok);
void orc_columns_mt(const uint8_t* data, const int64_t* off, int64_t n, int32_t* ref, int32_t* pos, int32_t* end,
                    uint8_t* mapq, uint16_t* flag, int32_t* nref, int32_t* npos, int32_t* tlen, int32_t* l_seq,
                    uint16_t* cl, uint16_t* cr, const int64_t* cig_off, uint32_t* cigar, const int64_t* seq_off,
                    uint8_t* seq, const int64_t* qual_off, uint8_t* qual, const char* tags, int n_tags,
                    uint8_t* t_type, int64_t* t_off, int32_t* t_len, int64_t* t_val, int n_threads);

#define CHECK(c)                                                  \
    do {                                                          \
        if (!(c)) {                                               \
            fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
            exit(2);                                              \
        }                                                         \
    } while (0)

static void* xmalloc(size_t n) {
    void* p = malloc(n ? n : 1);
    CHECK(p);
    return p;
}

static uint64_t rng = 88172645463325252ull;
static uint64_t next_rand(void) {
    rng ^= rng << 13, rng ^= rng >> 7, rng ^= rng << 17;
    return rng;
}

/* the BGZF file of one synthetic recipe: header blocks, record blocks (mixed levels), EOF; codec 0
   zlib, 1 libdeflate */
static uint8_t* make_file(int kind, int64_t n_rec, int64_t* size, int64_t* header_bytes, int n_threads, int codec) {
    void* g = cbs_open(kind, 7 + kind, n_rec);
    CHECK(g);
    uint8_t* raw = xmalloc(1 << 24);
    int64_t h = cbs_header(g, raw, 1 << 16);
    CHECK(h > 0);
    int64_t got = 0;
    int64_t r = cbs_records(g, raw + h, (1 << 24) - h, n_rec, &got);
    CHECK(r > 0 && got == n_rec);
    cbs_close(g);
    int64_t total = h + r, nb = (total + 65279) / 65280;
    int64_t* off = xmalloc(8 * nb);
    int32_t* len = xmalloc(4 * nb);
    int8_t* lvl = xmalloc(nb);
    uint8_t* fx = xmalloc(nb);
    for (int64_t i = 0; i < nb; i++) {
        off[i] = i * 65280;
        len[i] = (int32_t)(total - off[i] < 65280 ? total - off[i] : 65280);
        lvl[i] = (int8_t)((int[]){0, 1, 6, 9}[i % 4]);
        fx[i] = (uint8_t)(i % 5 == 1);
    }
    uint8_t* slots = xmalloc((size_t)nb * 65536);
    int32_t* bs = xmalloc(4 * nb);
    CHECK(cbs_bgzf_blocks_codec(raw, off, len, nb, lvl, fx, slots, 65536, bs, n_threads, codec) == 0);
    int64_t out = 0;
    for (int64_t i = 0; i < nb; i++) out += bs[i];
    static const uint8_t eof_block[28] = {0x1f, 0x8b, 8, 4, 0, 0, 0, 0, 0, 0xff, 6, 0, 'B', 'C', 2, 0, 0x1b, 0,
                                          3,    0,    0, 0, 0, 0, 0, 0, 0, 0};
    uint8_t* file = xmalloc(out + 28);
    int64_t p = 0;
    for (int64_t i = 0; i < nb; i++) {
        memcpy(file + p, slots + (size_t)i * 65536, bs[i]);
        p += bs[i];
    }
    memcpy(file + p, eof_block, 28);
    *size = p + 28;
    *header_bytes = h;
    free(raw), free(off), free(len), free(lvl), free(fx), free(slots), free(bs);
    return file;
}

/* scan and inflate; returns the inflated bytes (or NULL when a block fails) */
static uint8_t* inflate_all(const uint8_t* file, int64_t size, int eof, int64_t* out_n, int n_threads) {
    int64_t cap = size / 26 + 2, res[4];
    int64_t* boff = xmalloc(8 * cap);
    int32_t *bsz = xmalloc(4 * cap), *isz = xmalloc(4 * cap);
    uint32_t* crc = xmalloc(4 * cap);
    orc_bgzf_scan(file, size, eof, boff, bsz, isz, crc, cap, res);
    uint8_t* dst = NULL;
    *out_n = -1;
    if (res[2] == 0) {
        int64_t nb = res[0], tot = 0;
        int64_t* ooff = xmalloc(8 * (nb + 1));
        for (int64_t i = 0; i < nb; i++) ooff[i] = tot, tot += isz[i];
        dst = xmalloc(tot);
        uint32_t* got = xmalloc(4 * (nb + 1));
        int32_t* st = xmalloc(4 * (nb + 1));
        orc_inflate(file, boff, bsz, isz, ooff, nb, dst, got, st, n_threads);
        int bad = 0;
        for (int64_t i = 0; i < nb; i++) bad |= st[i] != 0;
        free(ooff), free(got), free(st);
        if (bad) {
            free(dst);
            dst = NULL;
        } else {
            *out_n = tot;
        }
    }
    free(boff), free(bsz), free(isz), free(crc);
    return dst;
}

static void columns(const uint8_t* d, int64_t n, int n_ref, int n_threads) {
    int64_t cap = n / 36 + 1, out[3];
    int64_t* off = xmalloc(8 * cap);
    orc_record_chain(d, n, n_ref, off, cap, out);
    int64_t m = out[0];
    int32_t *ncig = xmalloc(4 * m), *lseq = xmalloc(4 * m);
    orc_var_sizes(d, off, m, ncig, lseq);
    int64_t *co = xmalloc(8 * (m + 1)), *so = xmalloc(8 * (m + 1)), *qo = xmalloc(8 * (m + 1));
    co[0] = so[0] = qo[0] = 0;
    for (int64_t i = 0; i < m; i++) {
        co[i + 1] = co[i] + ncig[i];
        so[i + 1] = so[i] + (lseq[i] + 1) / 2;
        qo[i + 1] = qo[i] + lseq[i];
    }
    int32_t *ref = xmalloc(4 * m), *pos = xmalloc(4 * m), *end = xmalloc(4 * m), *nref = xmalloc(4 * m);
    int32_t *npos = xmalloc(4 * m), *tlen = xmalloc(4 * m), *ls = xmalloc(4 * m);
    uint8_t* mapq = xmalloc(m);
    uint16_t *flag = xmalloc(2 * m), *cl = xmalloc(2 * m), *cr = xmalloc(2 * m);
    uint32_t* cigar = xmalloc(4 * co[m]);
    uint8_t *seq = xmalloc(so[m]), *qual = xmalloc(qo[m]);
    uint8_t* tt = xmalloc(2 * m);
    int64_t *to = xmalloc(16 * m), *tv = xmalloc(16 * m);
    int32_t* tl = xmalloc(8 * m);
    orc_columns_mt(d, off, m, ref, pos, end, mapq, flag, nref, npos, tlen, ls, cl, cr, co, cigar, so, seq, qo, qual,
                   "CBUB", 2, tt, to, tl, tv, n_threads);
    orc_find_tag(d, off, m, 'x', 'f', tt, to, tl, tv);
    for (int64_t i = 0; i < m; i++) {
        int ok = 0;
        (void)orc_fnv1a(d + off[i] + 36, d[off[i] + 12] - 1);
        if (tt[i] == 'Z') (void)orc_pack_acgt(d + to[i], tl[i], &ok);
    }
    (void)orc_record_candidates(d, n < 200000 ? n : 200000, n_ref);
    free(off), free(ncig), free(lseq), free(co), free(so), free(qo), free(ref), free(pos), free(end), free(nref);
    free(npos), free(tlen), free(ls), free(mapq), free(flag), free(cl), free(cr), free(cigar), free(seq), free(qual);
    free(tt), free(to), free(tv), free(tl);
}

int main(void) {
    for (int run = 0; run < 4; run++) {
        const int kind = run % 2, codec = run / 2;
        if (codec == 1 && !cbs_have_libdeflate()) continue;
        int64_t size, hdr, n;
        uint8_t* file = make_file(kind, 20000, &size, &hdr, 4, codec);
        uint8_t* d = inflate_all(file, size, 1, &n, 4);
        CHECK(d && n > hdr);
        columns(d + hdr, n - hdr, 25, 4);
        columns(d + hdr, (n - hdr) * 2 / 3, 25, 1); /* a cut: the chain stops at an incomplete record */
        /* records with corrupted bytes: the chain stops where the test fails */
        uint8_t* c = xmalloc(n);
        for (int k = 0; k < 40; k++) {
            memcpy(c, d, n);
            for (int j = 0; j < 8; j++) c[hdr + next_rand() % (n - hdr)] ^= (uint8_t)(1 + next_rand() % 255);
            int64_t cap = (n - hdr) / 36 + 1, out[3];
            int64_t* off = xmalloc(8 * cap);
            orc_record_chain(c + hdr, n - hdr, 25, off, cap, out);
            free(off);
        }
        free(c);
        /* compressed bytes flipped, cut short, or random */
        uint8_t* f2 = xmalloc(size);
        for (int k = 0; k < 60; k++) {
            memcpy(f2, file, size);
            for (int j = 0; j < 4; j++) f2[next_rand() % size] ^= (uint8_t)(1 + next_rand() % 255);
            int64_t m;
            uint8_t* x = inflate_all(f2, size - (k % 3 == 0 ? (int64_t)(next_rand() % 5000) : 0), k & 1, &m, 2);
            free(x);
        }
        for (int64_t i = 0; i < size; i++) f2[i] = (uint8_t)next_rand();
        int64_t m;
        free(inflate_all(f2, size < 300000 ? size : 300000, 1, &m, 1));
        (void)orc_record_candidates(f2, size < 300000 ? size : 300000, 25);
        free(f2);
        free(d);
        free(file);
    }
    puts("sanitized run ok");
    return 0;
}
