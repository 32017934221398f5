/*
 * ref_render_shim.c -- TEST INFRASTRUCTURE ONLY: flat-argument entry points
 * into the reference's own renderer and one-shot builder, compiled with the
 * reference sources (oracle/Makefile, target ref) into
 * oracle/_ref/libmcref_render.so.  The parity tests call these to pin the
 * oracle's quad render and LOD decimation.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "matter_compressor.h"

/* mc_render_quad_lod over levels given as dense rasters.  blocked != 0 reads
 * every level through the reference's block path (dense pointer cleared, the
 * dense source's block() synthesizes 16^3 blocks) instead of its dense fast
 * path. */
int refshim_render_quad_lod(int nlods, const uint8_t *const *vols, const int *dims, const float *grid, int gw,
                            int gh, float x0, float y0, float step, int w, int h, int filter, int comp, float t0,
                            float t1, float dt, int blocked, uint8_t *out, int nthreads) {
    mc_sample_lods ls;
    memset(&ls, 0, sizeof ls);
    if (nlods < 1 || nlods > 8) return -1;
    for (int l = 0; l < nlods; ++l) {
        ls.lods[l] = mc_sample_src_dense(vols[l], dims[3 * l], dims[3 * l + 1], dims[3 * l + 2]);
        if (blocked) ls.lods[l].dense = NULL;
    }
    ls.nlods = nlods;
    mc_quad q;
    memset(&q, 0, sizeof q);
    q.grid = grid;
    q.gw = gw;
    q.gh = gh;
    mc_render_params p;
    memset(&p, 0, sizeof p);
    p.filter = (mc_filter)filter;
    p.comp = (mc_comp)comp;
    p.t0 = t0;
    p.t1 = t1;
    p.dt = dt;
    return mc_render_quad_lod(&ls, &q, x0, y0, step, w, h, &p, out, nthreads);
}

int refshim_pick_lod(int nlods, float vpp) {
    mc_sample_lods ls;
    memset(&ls, 0, sizeof ls);
    static const uint8_t one = 0;
    for (int l = 0; l < nlods && l < 8; ++l) ls.lods[l] = mc_sample_src_dense(&one, 1, 1, 1);
    ls.nlods = nlods;
    return mc_render_pick_lod(&ls, vpp);
}

float refshim_quad_spacing(const float *grid, int gw, int gh) {
    mc_quad q;
    memset(&q, 0, sizeof q);
    q.grid = grid;
    q.gw = gw;
    q.gh = gh;
    return mc_quad_spacing(&q);
}

typedef struct {
    const uint8_t *v;
    int nz, ny, nx;
} shim_vol;

static uint8_t shim_voxel(void *ud, int x, int y, int z) {
    const shim_vol *s = ud;
    return s->v[((size_t)z * s->ny + y) * s->nx + x];
}

/* mc_build of a dense volume at quality q; returns the archive bytes
 * (free with refshim_free) and their length. */
uint8_t *refshim_build(const uint8_t *vol, int nz, int ny, int nx, float q, size_t *out_len) {
    shim_vol s = {vol, nz, ny, nx};
    mc_build_opts o;
    memset(&o, 0, sizeof o);
    o.nx = nx;
    o.ny = ny;
    o.nz = nz;
    o.quality = q;
    return mc_build(shim_voxel, &s, &o, out_len);
}

void refshim_free(void *p) { free(p); }
