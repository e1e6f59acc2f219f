/*
 * cs_oracle.c -- serial CPU restatement of cell_simulation's RDME step (TEST INFRASTRUCTURE ONLY).
 *
 * Parity checker for the HIP path; never linked into the product. Each function cites the reference
 * lines it restates. Compile with -O2 -ffp-contract=off (no fast-math): every f32 operation must round
 * exactly as written, in the order written.
 */
#include "cs_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* random.wgsl:14-20. The LCG line 17 is dead code in the reference; the output hashes the input. */
uint32_t cso_pcg(uint32_t state)
{
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

/* random.wgsl:32-45 */
uint32_t cso_xorshift(uint32_t s)
{
    s ^= s << 13u;
    s ^= s >> 17u;
    s ^= s << 5u;
    return s;
}

/* random.wgsl:27-29: f32(u) * bitcast<f32>(0x2F800000) = f32(u) * 2^-32. */
float cso_uint_to_float(uint32_t u)
{
    float f = (float)u;          /* round-to-nearest-even conversion */
    return f * 2.3283064365386963e-10f;
}

/* random.wgsl:47-49 */
float cso_uniform_float(uint32_t state)
{
    return cso_uint_to_float(cso_xorshift(state));
}

float cso_expf(float x)
{
    return (float)exp((double)x);
}

static inline uint64_t nvox(const cso_model *m)
{
    return (uint64_t)m->res[0] * m->res[1] * m->res[2];
}

/* functions.wgsl:48-51 */
static inline uint32_t vidx(const cso_model *m, uint32_t x, uint32_t y, uint32_t z)
{
    return z + y * m->res[2] + x * m->res[2] * m->res[1];
}

/* rdme.wgsl:132-136: (D * tau) / (lambda * lambda), f32, D indexed [src][dst][sp]. */
static inline float prob_value(const cso_model *m, uint32_t src_reg, uint32_t dst_reg, uint32_t sp)
{
    uint32_t n1 = m->ns + 1u;
    float val = m->diffusion[sp + n1 * dst_reg + n1 * m->nreg * src_reg];
    float num = val * m->tau;
    float den = m->lambda * m->lambda;
    return num / den;
}

/* rdme.wgsl:138-191: cumulative hop probabilities in the order -x,+x,-y,+y,-z,+z; a missing
 * neighbour repeats the previous entry; p[6] = 1. */
void cso_probability_vector(const cso_model *m, uint32_t x, uint32_t y, uint32_t z, uint32_t sp,
                            float p[7])
{
    uint32_t src = m->regions[vidx(m, x, y, z)];
    if (x > 0u)
        p[0] = prob_value(m, src, m->regions[vidx(m, x - 1u, y, z)], sp);
    else
        p[0] = 0.0f;
    if (x < m->res[0] - 1u)
        p[1] = p[0] + prob_value(m, src, m->regions[vidx(m, x + 1u, y, z)], sp);
    else
        p[1] = p[0];
    if (y > 0u)
        p[2] = p[1] + prob_value(m, src, m->regions[vidx(m, x, y - 1u, z)], sp);
    else
        p[2] = p[1];
    if (y < m->res[1] - 1u)
        p[3] = p[2] + prob_value(m, src, m->regions[vidx(m, x, y + 1u, z)], sp);
    else
        p[3] = p[2];
    if (z > 0u)
        p[4] = p[3] + prob_value(m, src, m->regions[vidx(m, x, y, z - 1u)], sp);
    else
        p[4] = p[3];
    if (z < m->res[2] - 1u)
        p[5] = p[4] + prob_value(m, src, m->regions[vidx(m, x, y, z + 1u)], sp);
    else
        p[5] = p[4];
    p[6] = 1.0f;
}

/* rdme.wgsl:25-42 writeLatticeSite: first-fit over slots 0..S INCLUSIVE of the destination voxel.
 * Slot S is the next voxel's slot 0 (hazard H2). An index past the end of the lattice counts as an
 * occupied slot (the reference's out-of-bounds atomic result is undefined). Returns the slot index
 * written or -1. */
static int write_lattice_site(const cso_model *m, uint32_t *lat, uint64_t base, uint32_t value)
{
    uint64_t total = nvox(m) * m->S;
    for (uint32_t j = 0; j <= m->S; ++j) {
        uint64_t idx = base + j;
        if (idx < total && lat[idx] == 0u) {
            lat[idx] = value;
            return (int)j;
        }
    }
    return -1;
}

/* rdme.wgsl:45-88 move_particle_site, executed without interleaving (canonical order). */
static void move_particle_site(const cso_model *m, cso_state *st, uint32_t sp, uint32_t v_src,
                               uint32_t v_dst, uint64_t i_part, cso_counters *cnt)
{
    uint32_t n1 = m->ns + 1u;
    uint32_t old = st->occupancy[v_dst];
    st->occupancy[v_dst] = old + 1u;                    /* atomicAdd, :61 */
    if (old <= m->S) {                                  /* :62, hazard H3 */
        int j = write_lattice_site(m, st->lattice, (uint64_t)v_dst * m->S, sp);
        if (j >= 0) {
            st->counts[(uint64_t)v_dst * n1 + sp] += 1;   /* :66 */
            st->occupancy[v_src] -= 1u;                 /* :69 */
            if (i_part < nvox(m) * m->S)
                st->lattice[i_part] = 0u;               /* :70 */
            st->counts[(uint64_t)v_src * n1 + sp] -= 1;   /* :71 */
            if (cnt) {
                cnt->moves++;
                if ((uint32_t)j == m->S)
                    cnt->alias_writes++;
            }
            return;
        }
    }
    st->occupancy[v_dst] -= 1u;                         /* :75 / :80 */
    if (cnt)
        cnt->failed_claims++;
}

/* rdme.wgsl:205-226 for the particle at lattice index i_part walked by voxel (x, y, z): its hop draw and, with
 * move_particle (:91-130), the destination voxel, or UINT32_MAX when it stays (direction 6 or a reflecting
 * boundary). */
static uint32_t hop_target(const cso_model *m, uint32_t x, uint32_t y, uint32_t z, uint32_t sp, uint32_t i_part,
                           uint32_t itime)
{
    const uint32_t R0 = m->res[0], R1 = m->res[1], R2 = m->res[2];
    uint32_t xmy = (y != 0u) ? (x % y) : m->mod_zero;  /* hazard H4 */
    float p[7];
    cso_probability_vector(m, x, y, z, sp, p);
    uint32_t state = cso_pcg(itime + xmy + z + i_part);
    float r = cso_uniform_float(state);
    int i = 0;
    while (i < 6 && r > p[i])
        ++i;
    uint32_t dx = x, dy = y, dz = z;
    switch (i) {
    case 0: if (x == 0u) return UINT32_MAX; dx = x - 1u; break;
    case 1: if (x == R0 - 1u) return UINT32_MAX; dx = x + 1u; break;
    case 2: if (y == 0u) return UINT32_MAX; dy = y - 1u; break;
    case 3: if (y == R1 - 1u) return UINT32_MAX; dy = y + 1u; break;
    case 4: if (z == 0u) return UINT32_MAX; dz = z - 1u; break;
    case 5: if (z == R2 - 1u) return UINT32_MAX; dz = z + 1u; break;
    default: return UINT32_MAX;
    }
    return vidx(m, dx, dy, dz);
}

/* rdme.wgsl:194-229 (with move_particle :91-130), all voxels in ascending order. */
void cso_phase_r(const cso_model *m, const uint32_t *lat_src, const uint32_t *occ_src,
                 cso_state *st, uint32_t itime, cso_counters *cnt)
{
    const uint32_t R0 = m->res[0], R1 = m->res[1], R2 = m->res[2];
    const uint64_t total = nvox(m) * m->S;
    for (uint32_t x = 0; x < R0; ++x)
        for (uint32_t y = 0; y < R1; ++y)
            for (uint32_t z = 0; z < R2; ++z) {
                uint32_t v = vidx(m, x, y, z);
                uint32_t idx_lattice = v * m->S;             /* u32, functions.wgsl:42-46 */
                uint32_t occ = occ_src[v];
                for (uint32_t i_part = idx_lattice; i_part < idx_lattice + occ; ++i_part) {
                    uint32_t sp = (i_part < total) ? lat_src[i_part] : 0u;
                    if (sp == 0u)
                        continue;
                    uint32_t d = hop_target(m, x, y, z, sp, i_part, itime);
                    if (d != UINT32_MAX)
                        move_particle_site(m, st, sp, v, d, i_part, cnt);
                }
            }
}

/* cme.wgsl:25-120 for one voxel. */
static void cme_voxel(const cso_model *m, cso_state *st, int32_t *stat, uint32_t x, uint32_t y, uint32_t z,
                      uint32_t itime, float *cumm, cso_counters *cnt)
{
    const uint32_t n1 = m->ns + 1u, S = m->S, nr = m->nr;
    const uint32_t v = vidx(m, x, y, z);
    int32_t *C = st->counts + (uint64_t)v * n1;
    C[0] = 1;                                                       /* :31 */

    cumm[0] = 0.0f;                                                 /* :36-47 */
    float total = 0.0f;
    for (uint32_t r = 1; r <= nr; ++r) {
        float k = m->rates[r];
        const int32_t *ix = m->reactants + 3u * r;
        int32_t prod = (int32_t)((uint32_t)C[ix[0]] * (uint32_t)C[ix[1]] * (uint32_t)C[ix[2]]);
        float a = k * (float)prod;
        total = total + a;
        cumm[r] = cumm[r - 1] + a;
    }
    for (uint32_t i = 0; i < nr; ++i)                               /* :50-52 */
        cumm[i] = cumm[i] / total;

    uint32_t *L = st->lattice + (uint64_t)v * S;
    uint32_t cell_copy[16];                                         /* :56-59 */
    for (uint32_t i = 0; i < S; ++i)
        cell_copy[i] = L[i];

    uint32_t state = cso_pcg(itime + x * y + z);                    /* :64-66 */
    float r1 = cso_uniform_float(state);
    float xt = total * m->tau;
    float e = cso_expf(-xt);
    float thr = 1.0f - e;
    /* exp(-0) = 1 exactly, so only a non-zero exponent can leave the outcome to an exp ulp */
    if (cnt && xt != 0.0f && fabsf(r1 - thr) <= 5.9604644775390625e-08f)
        cnt->near_threshold++;
    if (!(r1 <= thr))
        return;
    if (cnt)
        cnt->fired++;

    float r2 = cso_uniform_float(state + 1u);                       /* :68-73 */
    uint32_t i = 0;
    while ((r2 > cumm[i]) & (i < nr))
        ++i;

    uint32_t j = 0;                                                 /* :76-119 */
    const int32_t *row = m->stoich + (uint64_t)i * n1;
    int32_t new_occ = (int32_t)st->occupancy[v];
    uint64_t res_idx = (uint64_t)v * S;                             /* hazard H6: lattice index */
    uint32_t res_sp = (m->reservoir && res_idx < nvox(m)) ? m->reservoir[res_idx] : 0u;
    for (uint32_t sp = 1; sp <= m->ns; ++sp) {
        if (res_sp == sp)
            continue;
        C[sp] += row[sp];
        new_occ += row[sp];
        stat[sp] += row[sp];
        int32_t cc = C[sp];
        while (cc > 0) {
            L[j] = sp;
            cc -= 1;
            j += 1;
            if (j >= S) {                                           /* rollback, :95-108 */
                for (uint32_t q = 0; q < S; ++q)
                    L[q] = cell_copy[q];
                for (uint32_t q = 1; q <= sp; ++q) {
                    C[q] -= row[q];
                    stat[q] -= row[q];
                }
                if (cnt)
                    cnt->rollbacks++;
                return;
            }
        }
    }
    st->occupancy[v] = (uint32_t)new_occ;                           /* :114 */
    while (j < S) {                                                 /* :116-119 */
        L[j] = 0u;
        j += 1;
    }
}

/* cme.wgsl over the dispatch of cme.rs:61-66. The x overhang of an odd R0 (hazard H5) is a no-op. */
void cso_phase_c(const cso_model *m, cso_state *st, uint32_t itime, cso_counters *cnt)
{
    float *cumm = (float *)malloc(sizeof(float) * (m->nr + 1u));
    for (uint32_t x = 0; x < m->res[0]; ++x)
        for (uint32_t y = 0; y < m->res[1]; ++y)
            for (uint32_t z = 0; z < m->res[2]; ++z)
                cme_voxel(m, st, st->stat, x, y, z, itime, cumm, cnt);
    free(cumm);
}

/* final_stage.wgsl:25-33: stable partition of non-zero slots to the front of each voxel. */
void cso_phase_f(const cso_model *m, cso_state *st)
{
    uint64_t n = nvox(m);
    for (uint64_t v = 0; v < n; ++v) {
        uint32_t *L = st->lattice + v * m->S;
        uint32_t j = 0;
        for (uint32_t i = 0; i < m->S; ++i) {
            if (L[i] != 0u) {
                uint32_t t = L[i];
                L[i] = L[j];
                L[j] = t;
                j += 1;
            }
        }
    }
}

/* simulation.rs:99-150: rdme -> cme -> final_stage; the dest->src copies make the next step's source
 * equal to this step's destination, so one in-place state plus a copy of the source suffices. */
int cso_step(const cso_model *m, cso_state *st, uint32_t itime, cso_counters *cnt)
{
    if (!m || !st || m->S == 0u || m->S > 16u || m->res[0] == 0u || m->res[1] == 0u || m->res[2] == 0u)
        return 1;
    uint64_t n = nvox(m);
    uint32_t *lat_src = (uint32_t *)malloc(sizeof(uint32_t) * n * m->S);
    uint32_t *occ_src = (uint32_t *)malloc(sizeof(uint32_t) * n);
    if (!lat_src || !occ_src) {
        free(lat_src);
        free(occ_src);
        return 2;
    }
    memcpy(lat_src, st->lattice, sizeof(uint32_t) * n * m->S);
    memcpy(occ_src, st->occupancy, sizeof(uint32_t) * n);
    cso_phase_r(m, lat_src, occ_src, st, itime, cnt);
    cso_phase_c(m, st, itime, cnt);
    cso_phase_f(m, st);
    free(lat_src);
    free(occ_src);
    return 0;
}

int cso_run(const cso_model *m, cso_state *st, const uint32_t *itime, uint32_t nsteps, cso_counters *cnt)
{
    for (uint32_t t = 0; t < nsteps; ++t) {
        int rc = cso_step(m, st, itime[t], cnt);
        if (rc)
            return rc;
    }
    return 0;
}

/* ---- the same step on several host threads (the CPU baseline) ------------------------------------------
 * The hop draws of phase R depend only on the source state, so the threads draw them for the first S walked
 * slots of every voxel (x-planes split among the threads) into a direction table; one thread then applies the
 * claims in the canonical order exactly as cso_phase_r does (a walk past slot S draws its extra hops there).
 * Phase C and phase F touch one voxel each and run on all threads, and so does the copy of the source state;
 * the global counts and the event counters are summed per thread. The result equals cso_step's bit for bit. */
typedef struct {
    const cso_model *m;
    cso_state *st;
    uint32_t *lat_src, *occ_src;
    uint8_t *dir;            /* [N*S] 0..5: hop direction; 6: stays or no particle */
    uint32_t *walkers;       /* this thread's voxels with a non-zero occupancy word, ascending */
    uint32_t n_walkers;
    uint32_t itime, x0, x1;
    int phase;               /* 0: source copy, 1: hop draws, 2: phase C, 3: phase F */
    int32_t *stat;           /* [ns+1] this thread's changes of the global counts */
    float *cumm;             /* [nr+1] */
    cso_counters cnt;
    char pad[64];            /* keeps the next thread's counters off this cache line */
} mt_job;

static uint8_t hop_dir(const cso_model *m, uint32_t x, uint32_t y, uint32_t z, uint32_t d)
{
    const uint32_t R1 = m->res[1], R2 = m->res[2];
    if (d == UINT32_MAX)
        return 6u;
    const uint32_t dx = d / (R1 * R2), dy = (d / R2) % R1, dz = d % R2;
    if (dx != x)
        return dx < x ? 0u : 1u;
    if (dy != y)
        return dy < y ? 2u : 3u;
    return dz < z ? 4u : 5u;
}

static void *mt_worker(void *arg)
{
    mt_job *j = (mt_job *)arg;
    const cso_model *m = j->m;
    const uint32_t R1 = m->res[1], R2 = m->res[2], S = m->S;
    const uint64_t total = nvox(m) * S;
    if (j->phase == 0) {  /* this slab's copy of the source state (the next voxel's slots come from the next slab) */
        const uint64_t v0 = (uint64_t)j->x0 * R1 * R2, v1 = (uint64_t)j->x1 * R1 * R2;
        memcpy(j->lat_src + v0 * S, j->st->lattice + v0 * S, sizeof(uint32_t) * (v1 - v0) * S);
        memcpy(j->occ_src + v0, j->st->occupancy + v0, sizeof(uint32_t) * (v1 - v0));
        return NULL;
    }
    if (j->phase == 1)
        j->n_walkers = 0;
    for (uint32_t x = j->x0; x < j->x1; ++x)
        for (uint32_t y = 0; y < R1; ++y)
            for (uint32_t z = 0; z < R2; ++z) {
                const uint32_t v = vidx(m, x, y, z);
                if (j->phase == 1) {
                    const uint32_t start = v * S, end = start + j->occ_src[v];
                    const uint32_t walk = end > start ? end - start : 0u;
                    if (j->occ_src[v] == 0u)
                        continue;
                    j->walkers[j->n_walkers++] = v;
                    for (uint32_t s = 0; s < S; ++s) {
                        const uint32_t ip = start + s;
                        const uint32_t sp = (s < walk && ip < total) ? j->lat_src[ip] : 0u;
                        j->dir[ip] = sp ? hop_dir(m, x, y, z, hop_target(m, x, y, z, sp, ip, j->itime)) : 6u;
                    }
                } else if (j->phase == 2) {
                    cme_voxel(m, j->st, j->stat, x, y, z, j->itime, j->cumm, &j->cnt);
                } else {
                    uint32_t *L = j->st->lattice + (uint64_t)v * S;
                    uint32_t k = 0;
                    for (uint32_t i = 0; i < S; ++i)
                        if (L[i] != 0u) {
                            uint32_t t = L[i];
                            L[i] = L[k];
                            L[k] = t;
                            k += 1;
                        }
                }
            }
    return NULL;
}

static void mt_phase(mt_job *jobs, uint32_t nt, int phase)
{
    pthread_t th[256];
    for (uint32_t t = 0; t < nt; ++t) {
        jobs[t].phase = phase;
        if (t > 0 && pthread_create(&th[t], NULL, mt_worker, &jobs[t]) != 0) {
            mt_worker(&jobs[t]);  /* no thread to spare: run it here */
            th[t] = 0;
        }
    }
    mt_worker(&jobs[0]);
    for (uint32_t t = 1; t < nt; ++t)
        if (th[t])
            pthread_join(th[t], NULL);
}

int cso_run_mt(const cso_model *m, cso_state *st, const uint32_t *itime, uint32_t nsteps, cso_counters *cnt,
               uint32_t nthreads)
{
    if (!m || !st || m->S == 0u || m->S > 16u || m->res[0] == 0u || m->res[1] == 0u || m->res[2] == 0u)
        return 1;
    const uint64_t n = nvox(m), total = n * m->S;
    uint32_t nt = nthreads < 1u ? 1u : (nthreads > 256u ? 256u : nthreads);
    if (nt > m->res[0])
        nt = m->res[0];
    uint32_t *lat_src = (uint32_t *)malloc(sizeof(uint32_t) * total);
    uint32_t *occ_src = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint8_t *dir = (uint8_t *)malloc(total);
    uint32_t *walkers = (uint32_t *)malloc(sizeof(uint32_t) * n);
    /* per-thread scratch on cache lines of its own (cme_voxel writes cumm for every voxel) */
    const size_t stat_stride = ((m->ns + 1u + 15u) / 16u + 1u) * 16u, cumm_stride = ((m->nr + 1u + 15u) / 16u + 1u) * 16u;
    int32_t *stats = (int32_t *)calloc((size_t)nt * stat_stride, sizeof(int32_t));
    float *cumm = (float *)malloc(sizeof(float) * (size_t)nt * cumm_stride);
    mt_job jobs[256];
    int rc = 0;
    if (!lat_src || !occ_src || !dir || !walkers || !stats || !cumm) {
        rc = 2;
        goto done;
    }
    for (uint32_t t = 0; t < nt; ++t) {
        memset(&jobs[t], 0, sizeof(mt_job));
        jobs[t].m = m;
        jobs[t].st = st;
        jobs[t].lat_src = lat_src;
        jobs[t].occ_src = occ_src;
        jobs[t].dir = dir;
        jobs[t].x0 = (uint32_t)((uint64_t)m->res[0] * t / nt);
        jobs[t].x1 = (uint32_t)((uint64_t)m->res[0] * (t + 1u) / nt);
        jobs[t].walkers = walkers + (uint64_t)jobs[t].x0 * m->res[1] * m->res[2];
        jobs[t].stat = stats + (size_t)t * stat_stride;
        jobs[t].cumm = cumm + (size_t)t * cumm_stride;
    }
    for (uint32_t k = 0; k < nsteps; ++k) {
        for (uint32_t t = 0; t < nt; ++t)
            jobs[t].itime = itime[k];
        mt_phase(jobs, nt, 0);
        mt_phase(jobs, nt, 1);
        /* phase R's claims in the canonical order (as cso_phase_r), over the voxels that walk any slot */
        const uint32_t R12 = m->res[1] * m->res[2];
        for (uint32_t t = 0; t < nt; ++t)
            for (uint32_t q = 0; q < jobs[t].n_walkers; ++q) {
                    const uint32_t v = jobs[t].walkers[q], start = v * m->S, occ = occ_src[v];
                    const uint32_t x = v / R12, y = (v / m->res[2]) % m->res[1], z = v % m->res[2];
                    for (uint32_t i_part = start; i_part < start + occ; ++i_part) {
                        const uint32_t sp = (i_part < total) ? lat_src[i_part] : 0u;
                        if (sp == 0u)
                            continue;
                        uint32_t d;
                        if (i_part - start < m->S) {
                            const uint8_t dd = dir[i_part];
                            if (dd >= 6u)
                                continue;
                            d = v;
                            switch (dd) {
                            case 0: d = v - m->res[1] * m->res[2]; break;
                            case 1: d = v + m->res[1] * m->res[2]; break;
                            case 2: d = v - m->res[2]; break;
                            case 3: d = v + m->res[2]; break;
                            case 4: d = v - 1u; break;
                            default: d = v + 1u; break;
                            }
                        } else {
                            d = hop_target(m, x, y, z, sp, i_part, itime[k]);
                            if (d == UINT32_MAX)
                                continue;
                        }
                        move_particle_site(m, st, sp, v, d, i_part, cnt);
                    }
            }
        mt_phase(jobs, nt, 2);
        mt_phase(jobs, nt, 3);
    }
    for (uint32_t t = 0; t < nt; ++t) {
        for (uint32_t sp = 0; sp <= m->ns; ++sp)
            st->stat[sp] += jobs[t].stat[sp];
        if (cnt) {
            cnt->fired += jobs[t].cnt.fired;
            cnt->rollbacks += jobs[t].cnt.rollbacks;
            cnt->near_threshold += jobs[t].cnt.near_threshold;
        }
    }
done:
    free(lat_src);
    free(occ_src);
    free(dir);
    free(walkers);
    free(stats);
    free(cumm);
    return rc;
}

/* ---- well-mixed (CME) phase: the build's own definition (cs_oracle.h), no reference counterpart ---- */

double cso_cme_h(const int32_t *row, const int32_t *counts)
{
    double h = 1.0;
    for (int j = 0; j < 3; ++j) {
        int32_t id = row[j];
        if (id < 0)
            continue;
        int seen = 0, m = 0;
        for (int q = 0; q < 3; ++q) {
            if (row[q] == id) {
                ++m;
                if (q < j)
                    seen = 1;
            }
        }
        if (seen)
            continue;
        double n = (double)counts[id];
        if (m == 1)
            h = h * n;
        else if (m == 2)
            h = h * (n * (n - 1.0) * 0.5);
        else
            h = h * (n * (n - 1.0) * (n - 2.0) / 6.0);
        if (!(h > 0.0))
            return 0.0;
    }
    return h;
}

uint64_t cso_cme_bits(uint32_t itime, uint32_t seed, uint32_t k)
{
    uint32_t a = cso_pcg(itime + cso_pcg(seed + 2u * k));
    uint32_t b = cso_pcg(itime + cso_pcg(seed + 2u * k + 1u));
    return (((uint64_t)a << 32) | b) >> 11;
}

static void kahan_add(float *s, float *c, float x)
{
    float y = x - *c;
    float t = *s + y;
    *c = (t - *s) - y;
    *s = t;
}

double cso_cme_total32(const cso_cme_model *c, const int32_t *counts)
{
    float s = 0.0f, k = 0.0f;
    for (uint32_t r = 0; r < c->ncr; ++r)
        kahan_add(&s, &k, c->rates[r] * (float)cso_cme_h(c->reactants + 3u * r, counts));
    return (double)s - (double)k;
}

uint32_t cso_cme_choice32(const cso_cme_model *c, const int32_t *counts, double u2)
{
    double target = u2 * cso_cme_total32(c, counts);
    float s = 0.0f, k = 0.0f;
    uint32_t chosen = c->ncr, last = c->ncr;
    for (uint32_t r = 0; r < c->ncr; ++r) {
        float ar = c->rates[r] * (float)cso_cme_h(c->reactants + 3u * r, counts);
        kahan_add(&s, &k, ar);
        if (ar > 0.0f) {
            last = r;
            if (chosen == c->ncr && target < (double)s - (double)k)
                chosen = r;
        }
    }
    return chosen == c->ncr ? last : chosen;
}

double cso_cme_total64(const cso_cme_model *c, const int32_t *counts)
{
    double s = 0.0;
    for (uint32_t r = 0; r < c->ncr; ++r)
        s = s + (double)c->rates[r] * cso_cme_h(c->reactants + 3u * r, counts);
    return s;
}

uint32_t cso_cme_choice64(const cso_cme_model *c, const int32_t *counts, double u2)
{
    double target = u2 * cso_cme_total64(c, counts), cum = 0.0;
    uint32_t chosen = c->ncr, last = c->ncr;
    for (uint32_t r = 0; r < c->ncr; ++r) {
        double ar = (double)c->rates[r] * cso_cme_h(c->reactants + 3u * r, counts);
        cum = cum + ar;
        if (ar > 0.0) {
            last = r;
            if (chosen == c->ncr && target < cum)
                chosen = r;
        }
    }
    return chosen == c->ncr ? last : chosen;
}

#define CSO_CME_MAX_EVENTS (1u << 20)

void cso_cme_phase(const cso_model *m, const cso_cme_model *c, cso_state *st, int32_t *counts, uint32_t itime,
                   cso_cme_counters *cnt)
{
    const double tau = (double)m->tau;
    double t = 0.0;
    uint32_t e = 0;
    for (;;) {
        double a0 = cso_cme_total32(c, counts);
        if (!(a0 > 0.0))
            break;
        if (e >= CSO_CME_MAX_EVENTS) {
            cnt->exceptions += 1;
            break;
        }
        double u1 = (double)(cso_cme_bits(itime, c->seed, 2u * e) + 1ull) * 0x1p-53;
        t = t + (-log(u1)) / a0;
        if (fabs(t - tau) <= tau * 1e-12)
            cnt->near += 1;
        if (t >= tau)
            break;
        double u2 = (double)cso_cme_bits(itime, c->seed, 2u * e + 1u) * 0x1p-53;
        uint32_t r = cso_cme_choice32(c, counts, u2);
        for (uint32_t k = 0; k < c->nc; ++k)
            counts[k] += c->delta[r * c->nc + k];
        int32_t rg = c->emit_region[r];
        for (uint32_t j = 0; j < 4u; ++j) {
            uint32_t sp = c->emit[4u * r + j];
            if (sp == 0u || rg < 0)
                continue;
            uint32_t n = c->region_len[rg];
            const uint32_t *list = c->region_voxels[rg];
            int placed = 0;
            for (uint32_t att = 0; att < 100u && n > 0u && !placed; ++att) {
                uint32_t u = cso_pcg(itime + cso_pcg((c->seed ^ 0x9E3779B9u) + ((e * 4u + j) << 7) + att));
                uint32_t v = list[(uint32_t)(((uint64_t)u * n) >> 32)];
                uint32_t *blk = st->lattice + (uint64_t)v * m->S;
                uint32_t occ = 0u;                   /* the voxel's particles (its slots are compacted) */
                while (occ < m->S && blk[occ] != 0u)
                    ++occ;
                if (occ < m->S) {
                    blk[occ] = sp;
                    st->occupancy[v] += 1u;
                    st->counts[(uint64_t)v * (m->ns + 1u) + sp] += 1;
                    st->stat[sp] += 1;
                    placed = 1;
                }
            }
            if (!placed)
                cnt->emit_fail += 1;
        }
        ++e;
    }
    cnt->events += e;
}

int cso_step_hybrid(const cso_model *m, const cso_cme_model *c, cso_state *st, int32_t *counts, uint32_t itime,
                    cso_counters *cnt, cso_cme_counters *ccnt)
{
    if (c && c->ncr > 0u)
        cso_cme_phase(m, c, st, counts, itime, ccnt);
    return cso_step(m, st, itime, cnt);
}

/* final_stage.wgsl:35-67 */
void cso_view(const cso_model *m, const cso_state *st, uint32_t view, uint32_t axis, uint32_t slice, float *out)
{
    const uint64_t total = nvox(m) * m->S;
    for (uint32_t x = 0; x < m->res[0]; ++x)
        for (uint32_t y = 0; y < m->res[1]; ++y)
            for (uint32_t z = 0; z < m->res[2]; ++z) {
                const uint32_t v = vidx(m, x, y, z);
                const uint32_t c[3] = {x, y, z};
                if (m->res[axis] - c[axis] < slice) {                        /* :35-38 */
                    out[v] = 0.0f;
                    continue;
                }
                if (view == 1u) {                                            /* :59-62 */
                    out[v] = (float)m->regions[v] / 255.0f;
                    continue;
                }
                if (view == 2u) {                                            /* :63-66 */
                    out[v] = (float)(m->reservoir ? m->reservoir[v] : 0u) / 255.0f;
                    continue;
                }
                const uint32_t occ = st->occupancy[v];                       /* :41-57 */
                uint32_t best = 0u, maxcount = 0u;
                if (occ <= 2u * m->S) {
                    const uint64_t base = (uint64_t)v * m->S;
                    for (uint32_t i = 0; i < occ; ++i) {
                        const uint32_t si = base + i < total ? st->lattice[base + i] : 0u;
                        uint32_t count = 0u;
                        for (uint32_t j = 0; j < occ; ++j)
                            count += (base + j < total ? st->lattice[base + j] : 0u) == si;
                        if (count > maxcount) {
                            maxcount = count;
                            best = si;
                        }
                    }
                }
                out[v] = (float)best / 255.0f;
            }
}
