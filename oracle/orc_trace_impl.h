/*
 * orc_trace_impl.h - voxelConeTracing.frag restated once, instantiated for
 * float (R = float, the reference's GLSL arithmetic) and double (FP64
 * reference used to state error). Included by orc_trace.c with R, FN() and
 * the M_* math macros defined. Test infrastructure only.
 */

/* voxelConeTracing.frag:78-135: fixed cone tables (The Tomorrow Children). */
static const R FN(DIRS16)[16][3] = {
    {0.57735, 0.57735, 0.57735},       {0.57735, -0.57735, -0.57735},
    {-0.57735, 0.57735, -0.57735},     {-0.57735, -0.57735, 0.57735},
    {-0.903007, -0.182696, -0.388844}, {-0.903007, 0.182696, 0.388844},
    {0.903007, -0.182696, 0.388844},   {0.903007, 0.182696, -0.388844},
    {-0.388844, -0.903007, -0.182696}, {0.388844, -0.903007, 0.182696},
    {0.388844, 0.903007, -0.182696},   {-0.388844, 0.903007, 0.182696},
    {-0.182696, -0.388844, -0.903007}, {0.182696, 0.388844, -0.903007},
    {-0.182696, 0.388844, 0.903007},   {0.182696, -0.388844, 0.903007}};

static const R FN(DIRS32)[32][3] = {
    {0.898904, 0.435512, 0.0479745},   {0.898904, -0.435512, -0.0479745},
    {0.898904, 0.0479745, -0.435512},  {0.898904, -0.0479745, 0.435512},
    {-0.898904, 0.435512, -0.0479745}, {-0.898904, -0.435512, 0.0479745},
    {-0.898904, 0.0479745, 0.435512},  {-0.898904, -0.0479745, -0.435512},
    {0.0479745, 0.898904, 0.435512},   {-0.0479745, 0.898904, -0.435512},
    {-0.435512, 0.898904, 0.0479745},  {0.435512, 0.898904, -0.0479745},
    {-0.0479745, -0.898904, 0.435512}, {0.0479745, -0.898904, -0.435512},
    {0.435512, -0.898904, 0.0479745},  {-0.435512, -0.898904, -0.0479745},
    {0.435512, 0.0479745, 0.898904},   {-0.435512, -0.0479745, 0.898904},
    {0.0479745, -0.435512, 0.898904},  {-0.0479745, 0.435512, 0.898904},
    {0.435512, -0.0479745, -0.898904}, {-0.435512, 0.0479745, -0.898904},
    {0.0479745, 0.435512, -0.898904},  {-0.0479745, -0.435512, -0.898904},
    {0.57735, 0.57735, 0.57735},       {0.57735, 0.57735, -0.57735},
    {0.57735, -0.57735, 0.57735},      {0.57735, -0.57735, -0.57735},
    {-0.57735, 0.57735, 0.57735},      {-0.57735, 0.57735, -0.57735},
    {-0.57735, -0.57735, 0.57735},     {-0.57735, -0.57735, -0.57735}};

static void FN(texel)(const orc_volume* v, long x, long y, long z, R out[4])
{
    long W = (long)(v->res + 2) * 6, H = (long)(v->res + 2) * v->levels, D = v->res + 2;
    /* VK_SAMPLER_ADDRESS_MODE_REPEAT (Voxelizer.cpp:279) */
    x %= W; if (x < 0) x += W;
    y %= H; if (y < 0) y += H;
    z %= D; if (z < 0) z += D;
    size_t i = ((size_t)z * (size_t)H + (size_t)y) * (size_t)W + (size_t)x;
    if (v->fmt == ORC_FMT_RGBA8) {
        uint32_t t = ((const uint32_t*)v->texels)[i];
        for (int k = 0; k < 4; ++k) out[k] = (R)((t >> (8 * k)) & 0xffu) / (R)255;
    } else {
        const uint16_t* t = (const uint16_t*)v->texels + 4u * i;
        for (int k = 0; k < 4; ++k) out[k] = (R)orc_half_to_float(t[k]);
    }
}

/* texture(sampler3D) with LINEAR filtering: u = s*W - 0.5, 8-tap trilinear. */
static void FN(tex3d)(const orc_volume* v, R s, R t, R r, R out[4])
{
    R W = (R)((v->res + 2) * 6), H = (R)((v->res + 2) * v->levels), D = (R)(v->res + 2);
    R u = s * W - (R)0.5, vv = t * H - (R)0.5, w = r * D - (R)0.5;
    R fu = M_FLOOR(u), fv = M_FLOOR(vv), fw = M_FLOOR(w);
    R a = u - fu, b = vv - fv, c = w - fw;
    long i0 = (long)fu, j0 = (long)fv, k0 = (long)fw;
    for (int k = 0; k < 4; ++k) out[k] = 0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                R wt = (dx ? a : (R)1 - a) * (dy ? b : (R)1 - b) * (dz ? c : (R)1 - c);
                R tx[4];
                FN(texel)(v, i0 + dx, j0 + dy, k0 + dz, tx);
                for (int k = 0; k < 4; ++k) out[k] += wt * tx[k];
            }
}

/* voxelConeTracing.frag:313-327 sampleClipmap */
static void FN(sample_clipmap)(const orc_volume* vol, const orc_trace_params* p, const R pos[3],
                               int level, const R face_off[3], const R weight[3], R out[4])
{
    R vs = (R)p->voxel_size * M_EXP2((R)level);
    R extent = vs * (R)p->volume_dimension;
    R sp[3];
    for (int k = 0; k < 3; ++k) {
        R q = pos[k] / extent;
        sp[k] = ((q - M_FLOOR(q)) * (R)p->volume_dimension + (R)1) / ((R)p->volume_dimension + (R)2);
    }
    sp[1] += (R)level;
    sp[1] /= (R)p->clip_levels;
    sp[0] /= (R)6;
    for (int k = 0; k < 4; ++k) out[k] = 0;
    for (int a = 0; a < 3; ++a) {
        R t[4];
        FN(tex3d)(vol, sp[0] + face_off[a], sp[1], sp[2], t);
        for (int k = 0; k < 4; ++k) out[k] += t[k] * weight[a];
    }
}

/* voxelConeTracing.frag:341-392 traceCone */
static void FN(trace_cone)(const orc_volume* vol, const orc_trace_params* p, const R start_in[3],
                           const R dir[3], R aperture, R max_dist, R start_level, R step_factor,
                           R out[4], uint64_t* steps)
{
    R result[4] = {0, 0, 0, 0};
    R cone_coef = (R)2 * M_TAN(aperture * (R)0.5);
    R cur_level = start_level;
    R voxel_size = (R)p->voxel_size * M_EXP2(cur_level);
    R start[3];
    for (int k = 0; k < 3; ++k) start[k] = start_in[k] + dir[k] * voxel_size * (R)p->trace_start_offset * (R)0.5;
    R step = 0;
    R diameter = M_FMAX(step * cone_coef, (R)p->voxel_size);
    R occlusion = 0;
    /* calculateVoxelFaceIndex (:296-303) */
    R face_off[3] = {(dir[0] > 0 ? (R)0 : (R)1) / (R)6, (dir[1] > 0 ? (R)2 : (R)3) / (R)6,
                     (dir[2] > 0 ? (R)4 : (R)5) / (R)6};
    R weight[3] = {dir[0] * dir[0], dir[1] * dir[1], dir[2] * dir[2]};
    R seg = voxel_size;
    R min_radius = (R)p->voxel_size * (R)p->volume_dimension * (R)0.5;
    const R max_level = (R)(p->clip_levels - 1);
    while (step < max_dist && occlusion < (R)1) {
        R pos[3];
        for (int k = 0; k < 3; ++k) pos[k] = start[k] + dir[k] * step;
        R dc[3] = {(R)p->volume_center[0] - pos[0], (R)p->volume_center[1] - pos[1],
                   (R)p->volume_center[2] - pos[2]};
        R dist = M_SQRT(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]);
        R min_level = M_CEIL(M_LOG2(dist / min_radius));
        cur_level = M_LOG2(diameter / (R)p->voxel_size);
        cur_level = M_FMIN(M_FMAX(M_FMAX(start_level, cur_level), min_level), max_level);
        /* sampleClipmapLinear (:329-339) */
        int lower = (int)M_FLOOR(cur_level), upper = (int)M_CEIL(cur_level);
        R lo[4], hi[4], s[4];
        FN(sample_clipmap)(vol, p, pos, lower, face_off, weight, lo);
        FN(sample_clipmap)(vol, p, pos, upper, face_off, weight, hi);
        R fl = cur_level - M_FLOOR(cur_level);
        for (int k = 0; k < 4; ++k) s[k] = lo[k] * ((R)1 - fl) + hi[k] * fl;
        voxel_size = (R)p->voxel_size * M_EXP2(cur_level);
        R correction = seg / voxel_size;
        R rad[3] = {s[0] * correction, s[1] * correction, s[2] * correction};
        /* 1 - a clamped at 0 (an FP filtered alpha can exceed 1 by an ulp;
         * the reference's filter weights sum to 1 exactly) */
        R opacity = (R)1 - M_POW(M_FMAX((R)1 - s[3], (R)0), correction);
        opacity = M_FMIN(M_FMAX(opacity, (R)0), (R)1);
        R blend = M_FMIN(M_FMAX((R)1 - result[3], (R)0), (R)1);
        result[0] += blend * rad[0];
        result[1] += blend * rad[1];
        result[2] += blend * rad[2];
        result[3] += blend * opacity;
        occlusion += ((R)1 - occlusion) * opacity / ((R)1 + (step + voxel_size) * (R)p->occlusion_decay);
        R prev = step;
        step += M_FMAX(diameter, (R)p->voxel_size) * step_factor;
        seg = step - prev;
        diameter = step * cone_coef;
        if (steps) ++*steps;
    }
    out[0] = result[0];
    out[1] = result[1];
    out[2] = result[2];
    out[3] = (R)1 - occlusion;
}

/* voxelConeTracing.frag:394-414 calcMinLevel */
static R FN(calc_min_level)(const orc_trace_params* p, const R pos[3])
{
    R dc[3] = {(R)p->volume_center[0] - pos[0], (R)p->volume_center[1] - pos[1],
               (R)p->volume_center[2] - pos[2]};
    R dist = M_SQRT(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]);
    R min_radius = (R)p->voxel_size * (R)p->volume_dimension * (R)0.5;
    R min_level = M_FMAX(M_LOG2(dist / min_radius), (R)0);
    R radius = min_radius * M_EXP2(M_CEIL(min_level));
    R f = dist / radius;
    const R ts = (R)0.5;
    R c = (R)1 / ((R)1 - ts);
    if (f > ts) return M_CEIL(min_level) + (f - ts) * c;
    return M_CEIL(min_level);
}

/* brdf.glsl:28-78 microfacetBRDF */
static void FN(brdf)(const R n[3], const R v[3], const R l[3], const R diffuse_color[3],
                     const R spec[3], R perceptual_roughness, R out[3])
{
    const R PI = (R)3.141592;
    R alpha = perceptual_roughness * perceptual_roughness;
    R reflectance = M_FMAX(M_FMAX(spec[0], spec[1]), spec[2]);
    R r90 = M_FMIN(M_FMAX(reflectance * (R)50, (R)0), (R)1);
    R h[3] = {l[0] + v[0], l[1] + v[1], l[2] + v[2]};
    R hl = M_SQRT(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    for (int k = 0; k < 3; ++k) h[k] /= hl;
    R ndl = M_FMIN(M_FMAX(n[0] * l[0] + n[1] * l[1] + n[2] * l[2], (R)0.001), (R)1);
    R ndv = M_FMIN(M_FMAX(M_FABS(n[0] * v[0] + n[1] * v[1] + n[2] * v[2]), (R)0.001), (R)1);
    R ndh = M_FMIN(M_FMAX(n[0] * h[0] + n[1] * h[1] + n[2] * h[2], (R)0), (R)1);
    R vdh = M_FMIN(M_FMAX(v[0] * h[0] + v[1] * h[1] + v[2] * h[2], (R)0), (R)1);
    R r = alpha;
    R att_l = (R)2 * ndl / (ndl + M_SQRT(r * r + ((R)1 - r * r) * (ndl * ndl)));
    R att_v = (R)2 * ndv / (ndv + M_SQRT(r * r + ((R)1 - r * r) * (ndv * ndv)));
    R G = att_l * att_v;
    R rsq = alpha * alpha;
    R f = (ndh * rsq - ndh) * ndh + (R)1;
    R D = rsq / (PI * f * f);
    R fres = M_POW(M_FMIN(M_FMAX((R)1 - vdh, (R)0), (R)1), (R)5);
    for (int k = 0; k < 3; ++k) {
        R F = spec[k] + (r90 - spec[k]) * fres;
        R diffuse_contrib = ((R)1 - F) * (diffuse_color[k] / PI);
        R spec_contrib = F * G * D / ((R)4 * ndl * ndv);
        out[k] = ndl * (diffuse_contrib + spec_contrib);
    }
}

static R FN(fetch_shadow)(const orc_shadow* sm, long i, long j)
{
    if (i < 0 || j < 0 || i >= sm->width || j >= sm->height) return 0;
    return (R)sm->depth[(size_t)j * (size_t)sm->width + (size_t)i];
}

static R FN(shadow_tap)(const orc_shadow* sm, R s, R t, int compare, R ref)
{
    R uu = s * (R)sm->width - (R)0.5, vv = t * (R)sm->height - (R)0.5;
    R fu = M_FLOOR(uu), fv = M_FLOOR(vv);
    R a = uu - fu, b = vv - fv;
    fu = M_FMIN(M_FMAX(fu, (R)-4), (R)sm->width + (R)4);
    fv = M_FMIN(M_FMAX(fv, (R)-4), (R)sm->height + (R)4);
    long i0 = (long)fu, j0 = (long)fv;
    R t4[4];
    long ii[4] = {i0, i0 + 1, i0, i0 + 1}, jj[4] = {j0, j0, j0 + 1, j0 + 1};
    for (int q = 0; q < 4; ++q) {
        R d = FN(fetch_shadow)(sm, ii[q], jj[q]);
        if (compare) {
            int in = ii[q] >= 0 && jj[q] >= 0 && ii[q] < sm->width && jj[q] < sm->height;
            d = (!in || ref <= d) ? (R)1 : (R)0;
        }
        t4[q] = d;
    }
    return ((R)1 - a) * ((R)1 - b) * t4[0] + a * ((R)1 - b) * t4[1] + ((R)1 - a) * b * t4[2] + a * b * t4[3];
}

/* shadow.glsl:28-36 */
static R FN(visibility)(const orc_shadow* sm, const orc_light* L, const R pos[3], uint32_t quirks)
{
    const float* V = L->view;
    const float* P = L->proj;
    R ls[3];
    for (int k = 0; k < 3; ++k)
        ls[k] = (R)V[0 + k] * pos[0] + (R)V[4 + k] * pos[1] + (R)V[8 + k] * pos[2] + (R)V[12 + k];
    int compare = !(quirks & ORC_QUIRK_SHADOW_RAW_DEPTH);
    R ref = 0;
    if (compare) ref = (R)P[2] * ls[0] + (R)P[6] * ls[1] + (R)P[10] * ls[2] + (R)P[14] - (R)0.001;
    ls[2] /= ((R)L->z_far - (R)L->z_near);
    R xy[2];
    for (int k = 0; k < 2; ++k) xy[k] = (R)P[0 + k] * ls[0] + (R)P[4 + k] * ls[1] + (R)P[12 + k];
    xy[0] = xy[0] * (R)0.5 + (R)0.5;
    xy[1] = xy[1] * (R)0.5 + (R)0.5;
    R sx = (R)1 / (R)sm->width, sy = (R)1 / (R)sm->height;
    R sum = 0;
    for (int yi = 0; yi < 4; ++yi)
        for (int xi = 0; xi < 4; ++xi)
            sum += FN(shadow_tap)(sm, xy[0] + ((R)xi - (R)1.5) * sx, xy[1] + ((R)yi - (R)1.5) * sy, compare, ref);
    return sum * (R)0.0625;
}

static void FN(normalize)(R v[3])
{
    R l = M_SQRT(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= l;
    v[1] /= l;
    v[2] /= l;
}

/* voxelConeTracing.frag:143-294 main, for rows [row_begin, row_end). */
static void FN(trace_rows)(const orc_volume* vol, const orc_gbuffer* g, const orc_trace_params* p,
                           const orc_shadow* sm, const orc_light* L, uint32_t quirks, int row_begin,
                           int row_end, R* dout, R* sout, uint64_t* samples)
{
    const int W = g->width, H = g->height;
    uint64_t total = 0;
    for (int py = row_begin; py < row_end; ++py)
        for (int px = 0; px < W; ++px) {
            size_t idx = (size_t)py * (size_t)W + (size_t)px;
            R* dO = dout + 4u * idx;
            R* sO = sout + 4u * idx;
            dO[0] = dO[1] = dO[2] = 0; dO[3] = 1;
            sO[0] = sO[1] = sO[2] = 0; sO[3] = 1;
            R depth = (R)g->depth[idx];
            if (depth == (R)1) continue;
            R tc[2] = {((R)px + (R)0.5) / (R)W, ((R)py + (R)0.5) / (R)H};
            R q[4] = {tc[0] * 2 - 1, tc[1] * 2 - 1, depth, 1};
            R wp[4];
            for (int k = 0; k < 4; ++k)
                wp[k] = (R)p->view_proj_inv[k] * q[0] + (R)p->view_proj_inv[4 + k] * q[1] +
                        (R)p->view_proj_inv[8 + k] * q[2] + (R)p->view_proj_inv[12 + k] * q[3];
            R world[3] = {wp[0] / wp[3], wp[1] / wp[3], wp[2] / wp[3]};
            R view[3] = {(R)p->eye_pos[0] - world[0], (R)p->eye_pos[1] - world[1], (R)p->eye_pos[2] - world[2]};
            FN(normalize)(view);
            uint32_t dt = g->diffuse[idx], st = g->specular[idx];
            R diffuse_color[3], spec_color[3];
            for (int k = 0; k < 3; ++k) {
                diffuse_color[k] = (R)((dt >> (8 * k)) & 0xffu) / (R)255;
                spec_color[k] = (R)((st >> (8 * k)) & 0xffu) / (R)255;
            }
            R rough = (R)((dt >> 24) & 0xffu) / (R)255;
            R metallic = (R)((st >> 24) & 0xffu) / (R)255;
            R normal[3], emission[3];
            for (int k = 0; k < 3; ++k) {
                normal[k] = (R)orc_half_to_float(g->normal[4u * idx + k]) * (R)2 - (R)1;
                emission[k] = (R)orc_half_to_float(g->emission[4u * idx + k]);
            }
            FN(normalize)(normal);
            int has_emission = emission[0] > 0 || emission[1] > 0 || emission[2] > 0;

            R min_level = FN(calc_min_level)(p, world);
            R vs = (R)p->voxel_size * M_EXP2(min_level);
            R start[3];
            for (int k = 0; k < 3; ++k) start[k] = world[k] + normal[k] * vs * (R)p->trace_start_offset;
            R indirect[4] = {0, 0, 0, 1};
            const int ncones = p->enable_32_cones ? 32 : 16;
            const R aperture = p->enable_32_cones ? (R)0.628319 : (R)0.872665;
            const R sf = M_FMAX((R)0.2, (R)p->min_trace_step_factor);
            for (int i = 0; i < ncones; ++i) {
                const R* d = p->enable_32_cones ? FN(DIRS32)[i] : FN(DIRS16)[i];
                R cos_theta = normal[0] * d[0] + normal[1] * d[1] + normal[2] * d[2];
                if (cos_theta < 0) continue;
                R c[4];
                FN(trace_cone)(vol, p, start, d, aperture, (R)30, min_level, sf, c, &total);
                for (int k = 0; k < 4; ++k) indirect[k] += c[k] * cos_theta;
            }
            for (int k = 0; k < 4; ++k) indirect[k] /= (R)ncones;
            indirect[3] *= (R)p->ambient_occlusion_factor;
            for (int k = 0; k < 3; ++k) indirect[k] *= diffuse_color[k] * (R)p->indirect_diffuse_intensity;

            R spec[3] = {0, 0, 0};
            if ((spec_color[0] > (R)1e-6 || spec_color[1] > (R)1e-6 || spec_color[2] > (R)1e-6) &&
                metallic > (R)1e-6) {
                /* reflect(-view, normal) */
                R I[3] = {-view[0], -view[1], -view[2]};
                R dni = normal[0] * I[0] + normal[1] * I[1] + normal[2] * I[2];
                R dir[3] = {I[0] - 2 * dni * normal[0], I[1] - 2 * dni * normal[1], I[2] - 2 * dni * normal[2]};
                R c[4];
                FN(trace_cone)(vol, p, start, dir, M_FMAX(rough, (R)0.05), (R)30, min_level,
                               (R)p->voxel_size, c, &total);
                for (int k = 0; k < 3; ++k) spec[k] = c[k] * spec_color[k] * (R)p->indirect_specular_intensity;
            }

            R direct[3] = {0, 0, 0};
            if (has_emission) {
                for (int k = 0; k < 3; ++k) direct[k] = emission[k];
            } else if (sm && L) {
                R l[3] = {-(R)L->direction[0], -(R)L->direction[1], -(R)L->direction[2]};
                FN(normalize)(l);
                R b[3];
                FN(brdf)(normal, view, l, diffuse_color, spec_color, rough, b);
                R vis = FN(visibility)(sm, L, world, quirks);
                for (int k = 0; k < 3; ++k) direct[k] = b[k] * vis;
            }

            switch (p->rendering_mode) {
            case 0: for (int k = 0; k < 3; ++k) dO[k] = diffuse_color[k]; break;
            case 1: for (int k = 0; k < 3; ++k) dO[k] = spec_color[k]; break;
            case 2: for (int k = 0; k < 3; ++k) dO[k] = normal[k] * (R)0.5 + (R)0.5; break;
            case 3: {
                static const R colors[7][4] = {{1, 0, 0, 1}, {0, 1, 0, 1}, {0, 0, 1, 1}, {1, 1, 0, 1},
                                               {0, 1, 1, 1}, {1, 0, 1, 1}, {1, 1, 1, 1}};
                int lower = (int)M_FLOOR(min_level);
                if (lower < 0) lower = 0;
                if (lower > 5) lower = 5;
                R fr = min_level - M_FLOOR(min_level);
                for (int k = 0; k < 4; ++k)
                    dO[k] = (colors[lower][k] * ((R)1 - fr) + colors[lower + 1][k] * fr) * (R)0.5;
                break;
            }
            case 4: for (int k = 0; k < 3; ++k) dO[k] = direct[k] * indirect[3]; break;
            case 5: for (int k = 0; k < 3; ++k) dO[k] = direct[k] * indirect[3] + indirect[k]; break;
            case 6: for (int k = 0; k < 3; ++k) sO[k] = spec[k]; break;
            case 7: for (int k = 0; k < 3; ++k) dO[k] = indirect[3]; break;
            case 8:
                for (int k = 0; k < 3; ++k) {
                    dO[k] = direct[k] * indirect[3] + indirect[k];
                    sO[k] = spec[k];
                }
                break;
            default: break;
            }
        }
    if (samples) *samples += total;
}
