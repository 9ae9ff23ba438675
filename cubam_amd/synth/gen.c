/*
 * Seeded synthetic BAM generator and parallel BGZF writer (host C, zlib).
 *
 * It produces the workloads the read path is measured on:
 *   kind 0  single-end 90 bp GEX-like reads: CIGAR 90M, or with p = 0.15 one
 *           N intron of 100..50,000 bp; mapq 0 (5 %) or 255; seq uniform ACGT
 *           with N at p = 0.001; qual uniform 2..41; tags CB:Z (16-mer + "-1"
 *           from a pool of 10,000), UB:Z (random 12-mer), xf:i.
 *   kind 1  paired-end 2 x 50 bp ATAC-like reads: proper pairs with insert
 *           size ~ lognormal(median 180, sigma 0.6), 2 % of pairs with one
 *           unmapped mate placed at its mate; CB:Z from a pool of 10,000
 *           barcodes drawn with Zipf(1.1) weights.
 * Both use 25 reference sequences of uniform 1..250 Mb and positions uniform
 * over the genome, emitted in coordinate order (sorted uniforms are drawn one
 * after the other with the descending-maximum method, so any prefix of a file
 * of n_total records is produced without generating the rest).
 *
 * Record bytes follow SAMv1 section 4.2 (block_size, refID, pos, l_read_name,
 * mapq, bin, n_cigar_op, flag, l_seq, next_refID, next_pos, tlen, read_name,
 * cigar, seq, qual, aux), the same layout the reference's test writer emits
 * (tests/bamwriter.py:74-113).  BGZF framing follows SAMv1 section 4.1 and
 * the reference's writer (tests/bamwriter.py:15-29): gzip member with FEXTRA
 * "BC" subfield, raw DEFLATE payload, CRC32 and ISIZE trailer, EOF block.
 * The payloads are compressed by zlib, or (codec 1, where the shared library
 * is present) by libdeflate at the same level: about three times faster on
 * these records, with compressed sizes within 1 % of zlib's.
 */
#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

/* ------------------------------------------------------------------ RNG */
typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static void rng_seed(rng_t* r, uint64_t seed) {
    for (int i = 0; i < 4; i++) r->s[i] = splitmix(&seed);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t rng_u64(rng_t* r) { /* xoshiro256** */
    uint64_t* s = r->s;
    uint64_t out = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return out;
}
static inline double rng_unif(rng_t* r) { return (double)(rng_u64(r) >> 11) * (1.0 / 9007199254740992.0); }
static inline uint32_t rng_below(rng_t* r, uint32_t n) { return (uint32_t)(((rng_u64(r) >> 32) * (uint64_t)n) >> 32); }
static inline double rng_normal(rng_t* r) {
    double u = rng_unif(r), v = rng_unif(r);
    if (u < 1e-300) u = 1e-300;
    return sqrt(-2.0 * log(u)) * cos(6.283185307179586 * v);
}

/* ------------------------------------------------------------------ record encoding */
static int reg2bin(int beg, int end) { /* SAMv1 5.3 */
    --end;
    if (beg >> 14 == end >> 14) return ((1 << 15) - 1) / 7 + (beg >> 14);
    if (beg >> 17 == end >> 17) return ((1 << 12) - 1) / 7 + (beg >> 17);
    if (beg >> 20 == end >> 20) return ((1 << 9) - 1) / 7 + (beg >> 20);
    if (beg >> 23 == end >> 23) return ((1 << 6) - 1) / 7 + (beg >> 23);
    if (beg >> 26 == end >> 26) return ((1 << 3) - 1) / 7 + (beg >> 26);
    return 0;
}
static inline void put32(uint8_t* p, uint32_t v) { p[0] = v; p[1] = v >> 8; p[2] = v >> 16; p[3] = v >> 24; }
static inline void put16(uint8_t* p, uint32_t v) { p[0] = v; p[1] = v >> 8; }

typedef struct {
    int32_t ref, pos, nref, npos, tlen;
    uint8_t mapq;
    uint16_t flag;
    const char* name; /* without NUL */
    int lname;
    const uint32_t* cigar;
    int ncig;
    const uint8_t* seq; /* 4-bit codes, one per base */
    const uint8_t* qual;
    int lseq;
    const uint8_t* aux;
    int laux;
} rec_t;

/* Encode one record at p; returns its size (block_size + 4). */
static int64_t put_record(uint8_t* p, const rec_t* r) {
    int rlen = 0;
    for (int i = 0; i < r->ncig; i++) {
        uint32_t op = r->cigar[i] & 15;
        if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) rlen += r->cigar[i] >> 4;
    }
    int bin = r->pos >= 0 ? reg2bin(r->pos, r->pos + (rlen ? rlen : 1)) : 4680;
    int64_t bs = 32 + (r->lname + 1) + 4LL * r->ncig + (r->lseq + 1) / 2 + r->lseq + r->laux;
    put32(p, (uint32_t)bs);
    put32(p + 4, (uint32_t)r->ref);
    put32(p + 8, (uint32_t)r->pos);
    p[12] = (uint8_t)(r->lname + 1);
    p[13] = r->mapq;
    put16(p + 14, (uint32_t)bin);
    put16(p + 16, (uint32_t)r->ncig);
    put16(p + 18, r->flag);
    put32(p + 20, (uint32_t)r->lseq);
    put32(p + 24, (uint32_t)r->nref);
    put32(p + 28, (uint32_t)r->npos);
    put32(p + 32, (uint32_t)r->tlen);
    uint8_t* q = p + 36;
    memcpy(q, r->name, r->lname);
    q[r->lname] = 0;
    q += r->lname + 1;
    for (int i = 0; i < r->ncig; i++, q += 4) put32(q, r->cigar[i]);
    for (int i = 0; i < r->lseq; i += 2) *q++ = (uint8_t)((r->seq[i] << 4) | (i + 1 < r->lseq ? r->seq[i + 1] : 0));
    memcpy(q, r->qual, r->lseq);
    q += r->lseq;
    memcpy(q, r->aux, r->laux);
    return bs + 4;
}

/* ------------------------------------------------------------------ generator state */
#define N_CONTIG 25
#define POOL 10000
#define MAXPEND 8192

typedef struct {
    int64_t gpos; /* genome coordinate: the heap key */
    uint64_t order; /* insertion order: ties keep it */
    int32_t pos, ref, tlen, npos, mapq, unmapped, seq;
    uint16_t flag;
    uint32_t cb;
    uint64_t name_id;
} pending_t;

typedef struct {
    int kind;
    rng_t rng;
    int64_t n_total; /* records (kind 0) or pairs (kind 1) of the virtual file */
    int64_t emitted; /* positions drawn so far */
    double maxv;     /* descending-maximum state */
    int64_t glen;    /* genome length */
    int64_t clen[N_CONTIG], cstart[N_CONTIG + 1];
    char cname[N_CONTIG][8];
    char pool[POOL][17];
    double zipf_cdf[POOL];
    pending_t pend[MAXPEND]; /* mates waiting for their position (kind 1), a binary heap on gpos */
    int npend;
    uint64_t name_counter;
    uint64_t push_counter;
    int64_t records_out;
    int64_t next_x;
    int have_next;
} gen_t;

static inline int before(const pending_t* a, const pending_t* b) {
    return a->gpos < b->gpos || (a->gpos == b->gpos && a->order < b->order);
}
static void heap_push(gen_t* g, pending_t x) {
    x.order = g->push_counter++;
    int i = g->npend++;
    while (i > 0) {
        int p = (i - 1) / 2;
        if (before(&g->pend[p], &x)) break;
        g->pend[i] = g->pend[p];
        i = p;
    }
    g->pend[i] = x;
}
static pending_t heap_pop(gen_t* g) {
    pending_t top = g->pend[0], x = g->pend[--g->npend];
    int i = 0;
    for (;;) {
        int c = 2 * i + 1;
        if (c >= g->npend) break;
        if (c + 1 < g->npend && before(&g->pend[c + 1], &g->pend[c])) c++;
        if (!before(&g->pend[c], &x)) break;
        g->pend[i] = g->pend[c];
        i = c;
    }
    if (g->npend) g->pend[i] = x;
    return top;
}

void* cbs_open(int kind, uint64_t seed, int64_t n_total) {
    gen_t* g = (gen_t*)calloc(1, sizeof(gen_t));
    if (!g) return NULL;
    g->kind = kind;
    rng_seed(&g->rng, seed * 0x100000001B3ULL + 17);
    g->n_total = kind == 1 ? (n_total + 1) / 2 : n_total;
    g->maxv = 1.0;
    int64_t s = 0;
    for (int c = 0; c < N_CONTIG; c++) {
        g->clen[c] = 1000000 + (int64_t)rng_below(&g->rng, 249000001);
        g->cstart[c] = s;
        s += g->clen[c];
        if (c < 22) snprintf(g->cname[c], 8, "chr%d", c + 1);
        else snprintf(g->cname[c], 8, "%s", c == 22 ? "chrX" : c == 23 ? "chrY" : "chrM");
    }
    g->cstart[N_CONTIG] = s;
    g->glen = s;
    for (int i = 0; i < POOL; i++) {
        for (int k = 0; k < 16; k++) g->pool[i][k] = "ACGT"[rng_below(&g->rng, 4)];
        g->pool[i][16] = 0;
    }
    double acc = 0;
    for (int i = 0; i < POOL; i++) g->zipf_cdf[i] = (acc += pow((double)(i + 1), -1.1));
    for (int i = 0; i < POOL; i++) g->zipf_cdf[i] /= acc;
    return g;
}

void cbs_close(void* h) { free(h); }

/* Uncompressed BAM header (magic, SAM text, references). Returns its size, or -1 when cap is too small. */
int64_t cbs_header(void* h, uint8_t* out, int64_t cap) {
    gen_t* g = (gen_t*)h;
    char text[4096];
    int t = snprintf(text, sizeof text, "@HD\tVN:1.6\tSO:coordinate\n");
    for (int c = 0; c < N_CONTIG; c++)
        t += snprintf(text + t, sizeof text - t, "@SQ\tSN:%s\tLN:%lld\n", g->cname[c], (long long)g->clen[c]);
    t += snprintf(text + t, sizeof text - t, "@PG\tID:synth\tPN:cubam_amd-synth\tVN:1\n");
    int64_t need = 12 + t;
    for (int c = 0; c < N_CONTIG; c++) need += 8 + (int64_t)strlen(g->cname[c]) + 1;
    if (need > cap) return -1;
    uint8_t* p = out;
    memcpy(p, "BAM\1", 4);
    put32(p + 4, (uint32_t)t);
    memcpy(p + 8, text, t);
    p += 8 + t;
    put32(p, N_CONTIG);
    p += 4;
    for (int c = 0; c < N_CONTIG; c++) {
        int l = (int)strlen(g->cname[c]) + 1;
        put32(p, (uint32_t)l);
        memcpy(p + 4, g->cname[c], l);
        put32(p + 4 + l, (uint32_t)g->clen[c]);
        p += 8 + l;
    }
    return p - out;
}

int64_t cbs_contig_length(void* h, int c) { return ((gen_t*)h)->clen[c]; }

/* The next sorted uniform position over the genome, or -1 when all are drawn. */
static int64_t next_gpos(gen_t* g) {
    if (g->emitted >= g->n_total) return -1;
    int64_t left = g->n_total - g->emitted;
    g->maxv *= pow(rng_unif(&g->rng), 1.0 / (double)left);
    g->emitted++;
    int64_t x = (int64_t)((1.0 - g->maxv) * (double)g->glen);
    if (x >= g->glen) x = g->glen - 1;
    if (x < 0) x = 0;
    return x;
}

static int contig_of(const gen_t* g, int64_t x) {
    int lo = 0, hi = N_CONTIG - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (g->cstart[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

static void make_name(char* buf, uint64_t id) {
    /* 28 characters, Illumina-like, unique per id */
    snprintf(buf, 29, "LH00345:12:H7:%06u:%07u", (unsigned)((id / 10000000ULL) % 1000000ULL),
             (unsigned)(id % 10000000ULL));
}

static void random_seq(rng_t* r, uint8_t* seq, int n) {
    static const uint8_t code[4] = {1, 2, 4, 8};
    for (int i = 0; i < n; i++) {
        uint64_t u = rng_u64(r);
        seq[i] = ((u >> 40) % 1000 == 0) ? 15 : code[u & 3];
    }
}

static void random_qual(rng_t* r, uint8_t* q, int n) {
    for (int i = 0; i < n; i++) q[i] = (uint8_t)(2 + rng_below(r, 40));
}

static int aux_z(uint8_t* p, const char* tag, const char* s, int n) {
    p[0] = tag[0]; p[1] = tag[1]; p[2] = 'Z';
    memcpy(p + 3, s, n);
    p[3 + n] = 0;
    return 4 + n;
}

/* kind 0: one single-end GEX-like record */
static int64_t gen_gex(gen_t* g, uint8_t* out, int64_t cap) {
    if (cap < 1024) return 0;
    int64_t x = next_gpos(g);
    if (x < 0) return -1;
    rng_t* r = &g->rng;
    int c = contig_of(g, x);
    int32_t pos = (int32_t)(x - g->cstart[c]);
    uint32_t cig[3];
    int ncig = 1;
    if (rng_unif(r) < 0.15) {
        int a = 1 + (int)rng_below(r, 89);
        int n = 100 + (int)rng_below(r, 49901);
        cig[0] = (uint32_t)a << 4; cig[1] = ((uint32_t)n << 4) | 3; cig[2] = (uint32_t)(90 - a) << 4;
        ncig = 3;
    } else {
        cig[0] = 90u << 4;
    }
    uint8_t seq[90], qual[90], aux[128];
    random_seq(r, seq, 90);
    random_qual(r, qual, 90);
    char bc[20], ub[13];
    memcpy(bc, g->pool[rng_below(r, POOL)], 16);
    bc[16] = '-'; bc[17] = '1';
    for (int k = 0; k < 12; k++) ub[k] = "ACGT"[rng_below(r, 4)];
    int la = aux_z(aux, "CB", bc, 18);
    la += aux_z(aux + la, "UB", ub, 12);
    static const int xf[4] = {25, 17, 0, 1};
    uint32_t u = rng_below(r, 100);
    aux[la] = 'x'; aux[la + 1] = 'f'; aux[la + 2] = 'i';
    put32(aux + la + 3, (uint32_t)xf[u < 70 ? 0 : u < 85 ? 1 : u < 95 ? 2 : 3]);
    la += 7;
    char name[32];
    make_name(name, g->name_counter++);
    rec_t rec = {c, pos, -1, -1, 0, (uint8_t)(rng_unif(r) < 0.05 ? 0 : 255), (uint16_t)(rng_below(r, 2) ? 16 : 0),
                 name, 28, cig, ncig, seq, qual, 90, aux, la};
    return put_record(out, &rec);
}

static uint32_t zipf_cell(gen_t* g) {
    double u = rng_unif(&g->rng);
    int lo = 0, hi = POOL - 1;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (g->zipf_cdf[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    return (uint32_t)lo;
}

static int64_t emit_atac(gen_t* g, const pending_t* m, uint8_t* out) {
    rng_t* r = &g->rng;
    uint8_t seq[50], qual[50], aux[32];
    random_seq(r, seq, 50);
    random_qual(r, qual, 50);
    char bc[20];
    memcpy(bc, g->pool[m->cb], 16);
    bc[16] = '-'; bc[17] = '1';
    int la = aux_z(aux, "CB", bc, 18);
    char name[32];
    make_name(name, m->name_id);
    uint32_t cig = 50u << 4;
    rec_t rec = {m->ref, m->pos, m->ref, m->npos, m->tlen, (uint8_t)m->mapq, m->flag, name, 28,
                 &cig, m->unmapped ? 0 : 1, seq, qual, 50, aux, la};
    g->records_out++;
    return put_record(out, &rec);
}

/* kind 1: fragments are drawn in coordinate order, one ahead; both mates wait in a heap on their
   genome position and are emitted once no fragment still to come can start before them. */
static void draw_fragment(gen_t* g, int64_t x) {
    rng_t* r = &g->rng;
    int c = contig_of(g, x);
    int32_t pos = (int32_t)(x - g->cstart[c]);
    int ins = (int)floor(exp(log(180.0) + 0.6 * rng_normal(r)) + 0.5);
    if (ins < 50) ins = 50;
    if (ins > 2000) ins = 2000;
    int32_t mpos = pos + ins - 50;
    uint32_t cb = zipf_cell(g);
    uint64_t id = g->name_counter++;
    int mapq = rng_unif(r) < 0.05 ? 0 : 60;
    uint16_t ra = rng_below(r, 2) ? 0x40 : 0x80, rb = ra ^ 0xC0;
    int unm = rng_unif(r) < 0.02 ? 1 + (int)rng_below(r, 2) : 0; /* 1: left mate unmapped, 2: right */
    pending_t a, b;
    memset(&a, 0, sizeof a);
    memset(&b, 0, sizeof b);
    a.ref = b.ref = c;
    a.cb = b.cb = cb;
    a.name_id = b.name_id = id;
    if (!unm) {
        a.pos = pos; a.npos = mpos; a.tlen = ins; a.mapq = mapq; a.flag = 0x1 | 0x2 | 0x20 | ra;
        b.pos = mpos; b.npos = pos; b.tlen = -ins; b.mapq = mapq; b.flag = 0x1 | 0x2 | 0x10 | rb;
        a.seq = 0; b.seq = 1;
    } else {
        /* the unmapped mate sits at its mate's position and follows it */
        pending_t* mapped = unm == 1 ? &b : &a;
        pending_t* um = unm == 1 ? &a : &b;
        a.pos = b.pos = pos; a.npos = b.npos = pos;
        mapped->mapq = mapq; mapped->flag = 0x1 | 0x8 | (mapped == &a ? ra : rb); mapped->seq = 0;
        um->mapq = 0; um->flag = 0x1 | 0x4 | (um == &a ? ra : rb); um->unmapped = 1; um->seq = 1;
    }
    a.gpos = g->cstart[c] + a.pos;
    b.gpos = g->cstart[c] + b.pos;
    heap_push(g, a);
    heap_push(g, b);
}

static int64_t gen_atac(gen_t* g, uint8_t* out, int64_t cap) {
    if (cap < 1024) return 0;
    for (;;) {
        if (!g->have_next && g->emitted < g->n_total) {
            g->next_x = next_gpos(g);
            g->have_next = 1;
        }
        if (g->npend > 0 && (!g->have_next || g->pend[0].gpos <= g->next_x)) {
            pending_t m = heap_pop(g);
            return emit_atac(g, &m, out);
        }
        if (!g->have_next) return -1;
        if (g->npend + 2 > MAXPEND) return -1; /* cannot happen with inserts <= 2,000 bp */
        g->have_next = 0;
        draw_fragment(g, g->next_x);
    }
}

/* Fill out with whole records (up to cap bytes, at most max_rec records). Returns bytes written;
   *n_rec receives the number of records; returns 0 with *n_rec = 0 at the end of the file. */
int64_t cbs_records(void* h, uint8_t* out, int64_t cap, int64_t max_rec, int64_t* n_rec) {
    gen_t* g = (gen_t*)h;
    int64_t used = 0, k = 0;
    while (k < max_rec) {
        int64_t n = g->kind == 1 ? gen_atac(g, out + used, cap - used) : gen_gex(g, out + used, cap - used);
        if (n <= 0) break;
        used += n;
        k++;
    }
    *n_rec = k;
    return used;
}

/* ------------------------------------------------------------------ parallel BGZF */
/* libdeflate's public compressor API, resolved at run time so that the generator needs neither its
   header nor the library itself */
typedef void* (*ld_alloc_fn)(int level);
typedef size_t (*ld_compress_fn)(void* c, const void* in, size_t n_in, void* out, size_t n_out);
typedef void (*ld_free_fn)(void* c);
static ld_alloc_fn ld_alloc;
static ld_compress_fn ld_compress;
static ld_free_fn ld_free;
static pthread_once_t ld_once = PTHREAD_ONCE_INIT;

static void ld_load(void) {
    void* h = dlopen("libdeflate.so.0", RTLD_NOW | RTLD_LOCAL);
    if (!h) return;
    ld_alloc = (ld_alloc_fn)dlsym(h, "libdeflate_alloc_compressor");
    ld_compress = (ld_compress_fn)dlsym(h, "libdeflate_deflate_compress");
    ld_free = (ld_free_fn)dlsym(h, "libdeflate_free_compressor");
    if (!ld_alloc || !ld_compress || !ld_free) ld_alloc = NULL;
}

/* 1 when libdeflate can be used (codec 1), else 0 */
int cbs_have_libdeflate(void) {
    pthread_once(&ld_once, ld_load);
    return ld_alloc != NULL;
}

typedef struct {
    const uint8_t* src;
    const int64_t* off;  /* payload start of each block in src */
    const int32_t* len;  /* payload bytes */
    const int8_t* level; /* zlib level per block */
    const uint8_t* fixed; /* Z_FIXED strategy per block (may be NULL) */
    uint8_t* dst;
    int64_t stride; /* bytes reserved per block in dst */
    int32_t* bsize; /* output: block size */
    int64_t n;
    int64_t next;
    pthread_mutex_t mu;
    int err;
    int codec; /* 0 zlib, 1 libdeflate (levels 1-9 without Z_FIXED; the rest by zlib) */
} job_t;

/* One BGZF block of src[0:len) at d (cap bytes): with ldc (a libdeflate compressor of the level)
   the payload is compressed by libdeflate, else by zlib. -2 when it does not fit a block. */
static int64_t frame_block(const uint8_t* src, int32_t len, int level, int fixed, void* ldc, uint8_t* d,
                           int64_t cap) {
    int64_t clen;
    if (ldc) {
        size_t got = ld_compress(ldc, src, (size_t)len, d + 18, (size_t)(cap - 26));
        if (got == 0) return -2;
        clen = (int64_t)got;
    } else {
        z_stream z;
        memset(&z, 0, sizeof z);
        if (deflateInit2(&z, level, Z_DEFLATED, -15, 8, fixed ? Z_FIXED : Z_DEFAULT_STRATEGY) != Z_OK) return -1;
        z.next_in = (Bytef*)src;
        z.avail_in = (uInt)len;
        z.next_out = d + 18;
        z.avail_out = (uInt)(cap - 26);
        int rc = deflate(&z, Z_FINISH);
        clen = (int64_t)z.total_out;
        deflateEnd(&z);
        if (rc != Z_STREAM_END) return -1;
    }
    int64_t bs = 18 + clen + 8;
    if (bs > 65536) return -2;
    static const uint8_t hdr[16] = {31, 139, 8, 4, 0, 0, 0, 0, 0, 255, 6, 0, 66, 67, 2, 0};
    memcpy(d, hdr, 16);
    put16(d + 16, (uint32_t)(bs - 1));
    put32(d + 18 + clen, (uint32_t)crc32(0L, src, (uInt)len));
    put32(d + 22 + clen, (uint32_t)len);
    return bs;
}

static void* worker(void* arg) {
    job_t* j = (job_t*)arg;
    void* ldc[10] = {0}; /* this thread's libdeflate compressors, by level */
    for (;;) {
        pthread_mutex_lock(&j->mu);
        int64_t i = j->next;
        j->next += 8;
        pthread_mutex_unlock(&j->mu);
        if (i >= j->n) break;
        int64_t e = i + 8 < j->n ? i + 8 : j->n;
        for (; i < e; i++) {
            int lvl = j->level ? j->level[i] : 6;
            int fx = j->fixed ? j->fixed[i] : 0;
            void* c = NULL;
            if (j->codec == 1 && lvl >= 1 && lvl <= 9 && !fx) {
                if (!ldc[lvl]) ldc[lvl] = ld_alloc(lvl);
                c = ldc[lvl];
                if (!c) j->err = 1;
            }
            int64_t bs = frame_block(j->src + j->off[i], j->len[i], lvl, fx, c, j->dst + i * j->stride, j->stride);
            if (bs == -2) /* did not fit: store it (level 0 never grows beyond len + 5 per 64 KiB) */
                bs = frame_block(j->src + j->off[i], j->len[i], 0, 0, NULL, j->dst + i * j->stride, j->stride);
            if (bs < 0) { j->err = 1; bs = 0; }
            j->bsize[i] = (int32_t)bs;
        }
    }
    for (int k = 0; k < 10; k++)
        if (ldc[k]) ld_free(ldc[k]);
    return NULL;
}

/* Compress n payloads (src[off[i]] .. +len[i], each <= 65,280 bytes) as BGZF blocks, in parallel.
   Block i is written at dst + i * stride (stride >= 65,536); bsize[i] receives its size. codec 1
   compresses with libdeflate (blocks at levels 1-9 without Z_FIXED). Returns 0, -1 on a compressor
   error, or -3 when codec 1 is asked for and libdeflate is absent. */
int cbs_bgzf_blocks_codec(const uint8_t* src, const int64_t* off, const int32_t* len, int64_t n,
                          const int8_t* level, const uint8_t* fixed, uint8_t* dst, int64_t stride, int32_t* bsize,
                          int n_threads, int codec) {
    if (codec == 1 && !cbs_have_libdeflate()) return -3;
    job_t j = {src, off, len, level, fixed, dst, stride, bsize, n, 0, PTHREAD_MUTEX_INITIALIZER, 0, codec};
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 256) n_threads = 256;
    pthread_t th[256];
    for (int t = 0; t < n_threads; t++) pthread_create(&th[t], NULL, worker, &j);
    for (int t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
    return j.err ? -1 : 0;
}

/* cbs_bgzf_blocks_codec with zlib */
int cbs_bgzf_blocks(const uint8_t* src, const int64_t* off, const int32_t* len, int64_t n, const int8_t* level,
                    const uint8_t* fixed, uint8_t* dst, int64_t stride, int32_t* bsize, int n_threads) {
    return cbs_bgzf_blocks_codec(src, off, len, n, level, fixed, dst, stride, bsize, n_threads, 0);
}
