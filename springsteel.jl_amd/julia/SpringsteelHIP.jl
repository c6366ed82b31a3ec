# SpringsteelHIP.jl -- ccall binding of libspringsteel_hip (MI355X transform
# library) for Springsteel.jl RL, RZ, RLZ, SL and SLZ grids.
#
# Loaded after Springsteel, it provides hip_spectralTransform!/hip_gridTransform!
# and their explicit-array forms with the reference's signatures and return
# values (src/transforms_cylindrical.jl:153-157, 254-257, 524-528, 654-657;
# src/transforms_cartesian.jl:627-631, 734-737; src/transforms_spherical.jl:
# 115-119, 213-216, 458-462, 582-585).  A plan is built once per grid and
# arithmetic form and cached by (objectid(grid), precision), like the
# reference's scratch registry (src/transform_scratch.jl:19, 167-185); the
# cache entry holds a weak reference to the grid's spectral array and is
# rebuilt when that array is no longer the grid's (an objectid reused after the
# grid was collected); _clear_hip_plan!(grid) releases the grid's plans.
# Every call checks the array sizes against the plan first (DimensionMismatch).
#
# physical / spectral may be host Arrays (the library copies them to the
# device and back) or device arrays whose pointer is passed with mem = DEVICE.
# A nonzero return code is rethrown as the exception the reference raises
# for the same condition.
module SpringsteelHIP

using Springsteel
using Springsteel: SpringsteelGrid, CubicBSpline, Chebyshev

const _RLGrid = Springsteel._RLGrid
const _RLZGrid = Springsteel._RLZGrid
const _RZGrid = Springsteel._2DCartesianRZ
const _SLGrid = Springsteel._SLGrid
const _SLZGrid = Springsteel._SLZGrid

const libssg = get(ENV, "SPRINGSTEEL_HIP_LIB",
                   joinpath(@__DIR__, "..", "springsteel_hip", "libspringsteel_hip.so"))

# return codes and enums of include/springsteel_hip.h
const SSG_OK, SSG_EINVAL, SSG_ENOTSUP, SSG_EPOSDEF, SSG_EDOMAIN, SSG_EBOUNDS = 0, -22, -95, -33, -34, -35
const SSG_RL, SSG_RZ, SSG_RLZ, SSG_SL, SSG_SLZ = Cint(1), Cint(2), Cint(3), Cint(4), Cint(5)
const SSG_MEM_HOST, SSG_MEM_DEVICE = Cint(0), Cint(1)
const SSG_FP64, SSG_FP32_REFINE = Cint(0), Cint(1)

# ssg_desc, field for field (isbits struct: C layout and padding)
struct SsgDesc
    geometry::Cint; num_cells::Cint; mubar::Cint; quadrature::Cint
    iMin::Float64; iMax::Float64
    nvars::Cint
    l_q::Ptr{Float64}
    bcl_kind::Ptr{Cint}; bcl_ab::Ptr{Float64}
    bcr_kind::Ptr{Cint}; bcr_ab::Ptr{Float64}
    kMin::Float64; kMax::Float64
    kDim::Cint; b_kDim::Cint
    bcb_kind::Ptr{Cint}; bct_kind::Ptr{Cint}
    max_wavenumber::Ptr{Cint}
    ring_kmax::Cint; patchOffsetL::Cint; precision::Cint
    filter_flags::Cint; wavenumber_ahat::Cint
end

# ssg_dims
struct SsgDims
    npts::Clong; nspec::Clong
    nvars::Cint; nslots::Cint
    iDim::Cint; b_iDim::Cint; jDim::Cint; kDim::Cint; b_kDim::Cint; kDim_wn::Cint
end

struct HIPPlan
    handle::Ptr{Cvoid}
    precision::Cint
    dims::SsgDims
    owner::WeakRef          # the spectral array of the grid the plan was built for
end

const _HIP_PLANS = Dict{Tuple{UInt, Cint}, HIPPlan}()

function _check_dims(p::HIPPlan, physical, spectral)
    d = p.dims
    np = Int(d.npts) * Int(d.nvars) * Int(d.nslots)
    ns = Int(d.nspec) * Int(d.nvars)
    physical === nothing || length(physical) == np ||
        throw(DimensionMismatch("physical has $(length(physical)) elements, the plan expects $np"))
    spectral === nothing || length(spectral) == ns ||
        throw(DimensionMismatch("spectral has $(length(spectral)) elements, the plan expects $ns"))
    return nothing
end

function _check(rc::Cint)
    rc == SSG_OK && return nothing
    msg = unsafe_string(ccall((:ssg_last_error, libssg), Cstring, ()))
    rc == SSG_EINVAL  && throw(ArgumentError(msg))
    rc == SSG_EPOSDEF && throw(ErrorException("PosDefException: " * msg))
    rc == SSG_EDOMAIN && throw(DomainError(rc, msg))
    rc == SSG_EBOUNDS && throw(BoundsError(msg))
    rc == SSG_ENOTSUP && throw(ErrorException("not supported by libspringsteel_hip: " * msg))
    throw(ErrorException("libspringsteel_hip error $rc: " * msg))
end

# Ooyama BC dict -> (kind, alpha, beta): GammaBC kinds of src/CubicBSpline.jl:290-294
function _spline_bc(bc::Dict)
    haskey(bc, "R0") && return (Cint(0), 0.0, 0.0)
    haskey(bc, "α1") && return (Cint(1), Float64(bc["α1"]), Float64(bc["β1"]))
    haskey(bc, "α2") && return (Cint(2), Float64(bc["α2"]), Float64(bc["β2"]))
    haskey(bc, "PERIODIC") && return (Cint(3), 0.0, 0.0)
    haskey(bc, "R3X") && return (Cint(5), 0.0, 0.0)
    haskey(bc, "R3") && return (Cint(4), 0.0, 0.0)
    throw(DomainError(bc, "Unrecognized spline BC dict keys"))
end

# Chebyshev BC dict -> kind: _bc_type, src/Chebyshev.jl:949-963
_cheb_bc(bc::Dict) = haskey(bc, "R0") ? Cint(0) : haskey(bc, "α0") ? Cint(1) : haskey(bc, "α1") ? Cint(2) : Cint(3)

_geometry(::_RLGrid) = SSG_RL
_geometry(::_RLZGrid) = SSG_RLZ
_geometry(::_RZGrid) = SSG_RZ
_geometry(::_SLGrid) = SSG_SL
_geometry(::_SLZGrid) = SSG_SLZ

# the ring cap createGrid applies (src/factory.jl:683, 731-734)
function _var_kmax(gp, key)
    get(gp.max_wavenumber, key, get(gp.max_wavenumber, "default", -1))
end
function _rlz_ring_kmax(gp)
    k = get(gp.max_wavenumber, "default", -1)
    for key in keys(gp.vars)
        k = get(gp.max_wavenumber, key, k)
    end
    return k
end

function _build_plan(grid::SpringsteelGrid, precision::Cint)
    gp = grid.params
    geom = _geometry(grid)
    nv = length(gp.vars)
    l_q = zeros(Float64, nv); bcl = zeros(Cint, nv); bcr = zeros(Cint, nv)
    bcl_ab = zeros(Float64, 2nv); bcr_ab = zeros(Float64, 2nv)
    bcb = zeros(Cint, nv); bct = zeros(Cint, nv); kmx = fill(Cint(-1), nv)
    for (key, v) in gp.vars
        sp = grid.ibasis.data[1, v].params            # the spline createGrid built for this variable
        l_q[v] = sp.l_q
        bcl[v], bcl_ab[2v-1], bcl_ab[2v] = _spline_bc(sp.BCL)
        bcr[v], bcr_ab[2v-1], bcr_ab[2v] = _spline_bc(sp.BCR)
        if geom != SSG_RL && geom != SSG_SL
            cp = grid.kbasis.data[v].params
            bcb[v] = _cheb_bc(cp.BCB); bct[v] = _cheb_bc(cp.BCT)
        end
        (geom == SSG_RL || geom == SSG_SL) && (kmx[v] = Cint(_var_kmax(gp, key)))
    end
    has_filters = false               # filter weights go through ssg_plan_set_filter (_hip_set_filter!)
    ahat = false                      # background coefficients go through hip_set_ahat! (ssg_plan_set_ahat)
    GC.@preserve l_q bcl bcr bcl_ab bcr_ab bcb bct kmx begin
        d = SsgDesc(geom, Cint(gp.num_cells), Cint(gp.mubar), Cint(gp.quadrature == :gauss ? 0 : 1),
                    Float64(gp.iMin), Float64(gp.iMax), Cint(nv),
                    pointer(l_q), pointer(bcl), pointer(bcl_ab), pointer(bcr), pointer(bcr_ab),
                    Float64(gp.kMin), Float64(gp.kMax), Cint(gp.kDim), Cint(gp.b_kDim),
                    pointer(bcb), pointer(bct), pointer(kmx),
                    Cint(geom == SSG_RLZ || geom == SSG_SLZ ? _rlz_ring_kmax(gp) : -1), Cint(gp.patchOffsetL), precision,
                    Cint(has_filters), Cint(ahat))
        out = Ref{Ptr{Cvoid}}(C_NULL)
        _check(ccall((:ssg_plan_create, libssg), Cint, (Ref{SsgDesc}, Ref{Ptr{Cvoid}}), d, out))
        dims = Ref{SsgDims}()
        _check(ccall((:ssg_plan_get_dims, libssg), Cint, (Ptr{Cvoid}, Ref{SsgDims}), out[], dims))
        return HIPPlan(out[], precision, dims[], WeakRef(grid.spectral))
    end
end

# precision: SSG_FP64 (the drop-in) or SSG_FP32_REFINE (physical in Float32, FP32
# synthesis, FP64 forward path, solve and evaluation; RL, RLZ, SL and SLZ grids)
function hip_plan(grid::SpringsteelGrid; precision::Cint = SSG_FP64)
    key = (objectid(grid), precision)
    p = get(_HIP_PLANS, key, nothing)
    if p !== nothing && p.owner.value !== grid.spectral
        # the objectid of a collected grid, reused: that plan is not this grid's
        ccall((:ssg_plan_destroy, libssg), Cint, (Ptr{Cvoid},), p.handle)
        delete!(_HIP_PLANS, key)
        p = nothing
    end
    if p === nothing
        p = _build_plan(grid, precision)
        _HIP_PLANS[key] = p
        _hip_set_filter!(grid, p)
    end
    return p
end

# applyFilter! (src/filtering.jl:262-428).  The weights come from the reference's
# own _get_filter / _filter_weight, so the library receives its values bit for bit;
# column v of each table holds variable v's weights (1.0 where its dict has no entry).
function _hip_filter_tables(grid::SpringsteelGrid)
    gp = grid.params
    K = gp.iDim + gp.patchOffsetL
    function table(d, n)
        isempty(d) && return nothing
        w = ones(Float64, n, length(gp.vars))
        for (name, v) in gp.vars
            f = Springsteel._get_filter(d, name)
            f === nothing && continue
            for k in 0:(n - 1)
                w[k + 1, v] = Springsteel._filter_weight(f, k)
            end
        end
        return w
    end
    return table(gp.fourier_filter, K + 1), table(gp.chebyshev_filter, gp.b_kDim)
end

function _hip_set_filter!(grid::SpringsteelGrid, p)
    fw, cw = _hip_filter_tables(grid)
    (fw === nothing && cw === nothing) && return nothing
    fp = fw === nothing ? Ptr{Float64}(C_NULL) : pointer(fw)
    cp = cw === nothing ? Ptr{Float64}(C_NULL) : pointer(cw)
    GC.@preserve fw cw _check(ccall((:ssg_plan_set_filter, libssg), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                                    p.handle, fp, cp))
    return nothing
end

function _clear_hip_plan!(grid::SpringsteelGrid)
    for prec in (SSG_FP64, SSG_FP32_REFINE)
        p = pop!(_HIP_PLANS, (objectid(grid), prec), nothing)
        p === nothing || ccall((:ssg_plan_destroy, libssg), Cint, (Ptr{Cvoid},), p.handle)
    end
    return nothing
end

const _HipGrid = Union{_RLGrid, _RLZGrid, _RZGrid, _SLGrid, _SLZGrid}
# regular / unstructured evaluation: SL / SLZ take the RL / RLZ paths, as the reference's
# do (src/transforms_spherical.jl:849-1131; src/interpolation.jl:1782-1800)
const _EvalGrid = Union{_RLGrid, _RLZGrid, _RZGrid, _SLGrid, _SLZGrid}
const _EvalRL = Union{_RLGrid, _SLGrid}
const _EvalRLZ = Union{_RLZGrid, _SLZGrid}

# explicit-array forms: spectralTransform(grid, physical, spectral) / gridTransform(grid, physical, spectral)
function hip_spectralTransform(grid::_HipGrid, physical::Array{Float64}, spectral::Array{Float64};
                               stream::Ptr{Cvoid} = C_NULL)
    p = hip_plan(grid)
    _check_dims(p, physical, spectral)
    GC.@preserve physical spectral _check(ccall((:ssg_spectral_transform, libssg), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}),
        p.handle, pointer(physical), pointer(spectral), SSG_MEM_HOST, stream))
    return spectral
end

# the FP32 form: physical in Float32 (same layout), spectral in Float64
function hip_spectralTransform(grid::_HipGrid, physical::Array{Float32}, spectral::Array{Float64};
                               stream::Ptr{Cvoid} = C_NULL)
    p = hip_plan(grid; precision = SSG_FP32_REFINE)
    _check_dims(p, physical, spectral)
    GC.@preserve physical spectral _check(ccall((:ssg_spectral_transform_f32, libssg), Cint,
        (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float64}, Cint, Ptr{Cvoid}),
        p.handle, pointer(physical), pointer(spectral), SSG_MEM_HOST, stream))
    return spectral
end

# R3X background coefficients.  The reference keeps them per spectral slot in the
# wavenumber registry (src/multipatch_registry.jl:23-60; RL / SL slot 2k is the
# cosine line 2k - 1 and slot 2k + 1 the sine line, src/transforms_cylindrical.jl:
# 280-325; RLZ / SLZ slots are the lines, 690-749) or on the spline objects
# (set_ahat_r3x!: RL / RLZ / SL / SLZ rows 1, 2, 3 serve the k = 0, cosine and sine
# lines, RZ row z its z mode).  The library takes one b_iDim vector per spectral line.
_has_r3x(grid) = any(Springsteel.CubicBSpline._has_r3x(grid.ibasis.data[1, v].params) for v in values(grid.params.vars))

function _hip_ahat_matrix(grid::_HipGrid)
    gp = grid.params
    nv = length(gp.vars)
    M = gp.b_iDim
    nlines = size(grid.spectral, 1) ÷ M
    A = zeros(Float64, size(grid.spectral, 1), nv)
    twod = grid isa _RLGrid || grid isa _SLGrid
    per_block = grid isa _RZGrid ? 1 : 1 + 2 * (gp.iDim + gp.patchOffsetL)
    for v in 1:nv, line in 0:(nlines - 1)
        rows = (line * M + 1):((line + 1) * M)
        if Springsteel._has_wavenumber_ahat(grid)
            slot = (twod && line > 0) ? line + 1 : line
            A[rows, v] .= Springsteel._get_wavenumber_ahat(grid, v, slot)
        elseif grid isa _RZGrid
            A[rows, v] .= grid.ibasis.data[line + 1, v].ahat
        else
            p = line % per_block
            row = p == 0 ? 1 : (isodd(p) ? 2 : 3)
            A[rows, v] .= grid.ibasis.data[row, v].ahat
        end
    end
    return A
end

# hands the grid's current background coefficients to the plan (R3X grids only)
function hip_set_ahat!(grid::_HipGrid, ahat::Union{Matrix{Float64}, Nothing} = _hip_ahat_matrix(grid);
                       precision::Cint = SSG_FP64)
    p = hip_plan(grid; precision = precision)
    _check_dims(p, nothing, ahat)
    if ahat === nothing
        _check(ccall((:ssg_plan_set_ahat, libssg), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint), p.handle, C_NULL, SSG_MEM_HOST))
    else
        GC.@preserve ahat _check(ccall((:ssg_plan_set_ahat, libssg), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint),
                                       p.handle, ahat, SSG_MEM_HOST))
    end
    return nothing
end

function hip_gridTransform(grid::_HipGrid, physical::Array{Float64}, spectral::Array{Float64};
                           stream::Ptr{Cvoid} = C_NULL)
    _has_r3x(grid) && hip_set_ahat!(grid)
    p = hip_plan(grid)
    _check_dims(p, physical, spectral)
    GC.@preserve physical spectral _check(ccall((:ssg_grid_transform, libssg), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}),
        p.handle, pointer(spectral), pointer(physical), SSG_MEM_HOST, stream))
    return physical
end

function hip_gridTransform(grid::_HipGrid, physical::Array{Float32}, spectral::Array{Float64};
                           stream::Ptr{Cvoid} = C_NULL)
    _has_r3x(grid) && hip_set_ahat!(grid; precision = SSG_FP32_REFINE)
    p = hip_plan(grid; precision = SSG_FP32_REFINE)
    _check_dims(p, physical, spectral)
    GC.@preserve physical spectral _check(ccall((:ssg_grid_transform_f32, libssg), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float32}, Cint, Ptr{Cvoid}),
        p.handle, pointer(spectral), pointer(physical), SSG_MEM_HOST, stream))
    return physical
end

function hip_applyFilter!(grid::_HipGrid; stream::Ptr{Cvoid} = C_NULL)
    p = hip_plan(grid)
    _check_dims(p, nothing, grid.spectral)
    GC.@preserve grid _check(ccall((:ssg_apply_filter, libssg), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint, Ptr{Cvoid}),
                                   p.handle, pointer(grid.spectral), SSG_MEM_HOST, stream))
    return nothing
end

# in-place forms with the reference's return values; spectralTransform! ends with
# applyFilter! (src/transforms_cylindrical.jl:153-157, 524-528)
function hip_spectralTransform!(grid::_HipGrid)
    hip_spectralTransform(grid, grid.physical, grid.spectral)
    hip_applyFilter!(grid)
    return grid.spectral
end
hip_gridTransform!(grid::_HipGrid) = hip_gridTransform(grid, grid.physical, grid.spectral)

# device-resident arrays (e.g. ROCArray): pass raw device pointers
function hip_spectralTransform_device(grid::_HipGrid, physical_ptr::Ptr{Float64}, spectral_ptr::Ptr{Float64};
                                      stream::Ptr{Cvoid} = C_NULL)
    _check(ccall((:ssg_spectral_transform, libssg), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}),
                 hip_plan(grid).handle, physical_ptr, spectral_ptr, SSG_MEM_DEVICE, stream))
end

function hip_gridTransform_device(grid::_HipGrid, spectral_ptr::Ptr{Float64}, physical_ptr::Ptr{Float64};
                                  stream::Ptr{Cvoid} = C_NULL)
    _check(ccall((:ssg_grid_transform, libssg), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}),
                 hip_plan(grid).handle, spectral_ptr, physical_ptr, SSG_MEM_DEVICE, stream))
end

# launch-shape tuner for one plan on its device arrays (overwrites them like a round trip);
# physical_ptr points to Float64, or Float32 for the FP32 form
function hip_tune!(grid::_HipGrid, physical_ptr::Ptr, spectral_ptr::Ptr{Float64};
                   stream::Ptr{Cvoid} = C_NULL, precision::Cint = SSG_FP64)
    _check(ccall((:ssg_plan_tune, libssg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid}),
                 hip_plan(grid; precision = precision).handle, physical_ptr, spectral_ptr, stream))
end

# tuning of a plan: (syn_ct, syn_nwq, ana_mr, si_seg, sa_mode), 0 = built-in choice
function hip_tuning(grid::_HipGrid)
    out = zeros(Cint, 5)
    _check(ccall((:ssg_plan_get_tuning, libssg), Cint, (Ptr{Cvoid}, Ptr{Cint}), hip_plan(grid).handle, out))
    return (syn_ct = out[1], syn_nwq = out[2], ana_mr = out[3], si_seg = out[4], sa_mode = out[5])
end

# pin a tuning, e.g. one saved from hip_tuning after hip_tune! in an earlier run
function hip_set_tuning!(grid::_HipGrid, t::NamedTuple)
    v = Cint[t.syn_ct, t.syn_nwq, t.ana_mr, t.si_seg, t.sa_mode]
    _check(ccall((:ssg_plan_set_tuning, libssg), Cint, (Ptr{Cvoid}, Ptr{Cint}), hip_plan(grid).handle, v))
end

# regularGridTransform at the tensor product of the points (RL: r x lambda,
# src/transforms_cylindrical.jl:929; RLZ: r x lambda x z, 1168; RZ: x x z,
# src/transforms_cartesian.jl:946): returns physical (npr, nvars, nslots), the
# last dimension fastest, as the reference's loops fill it
function _hip_regular(grid::_EvalGrid, i_pts, j_pts, k_pts)
    p = hip_plan(grid)
    nv = length(grid.params.vars)
    ns = grid isa _EvalRLZ ? 7 : 5
    iv = collect(Float64, i_pts)
    jv = j_pts === nothing ? Float64[] : collect(Float64, j_pts)
    kv = k_pts === nothing ? Float64[] : collect(Float64, k_pts)
    npr = length(iv) * max(length(jv), 1) * max(length(kv), 1)
    physical = zeros(Float64, npr, nv, ns)
    GC.@preserve iv jv kv physical _check(ccall((:ssg_regular_transform, libssg), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Ptr{Float64}, Cint,
         Ptr{Cvoid}),
        p.handle, pointer(grid.spectral), iv, Cint(length(iv)), jv, Cint(length(jv)), kv, Cint(length(kv)),
        pointer(physical), SSG_MEM_HOST, C_NULL))
    return physical
end

hip_regularGridTransform(grid::_EvalRL, r_pts::AbstractVector{Float64}, λ_pts::AbstractVector{Float64}) =
    _hip_regular(grid, r_pts, λ_pts, nothing)
hip_regularGridTransform(grid::_EvalRLZ, r_pts::AbstractVector{Float64}, λ_pts::AbstractVector{Float64},
                         z_pts::AbstractVector{Float64}) = _hip_regular(grid, r_pts, λ_pts, z_pts)
hip_regularGridTransform(grid::_RZGrid, x_pts::AbstractVector{Float64}, z_pts::AbstractVector{Float64}) =
    _hip_regular(grid, x_pts, nothing, z_pts)

# points-matrix forms, reduced to vectors as the reference does
# (src/transforms_cylindrical.jl:1014-1020; RZ / RLZ: sorted unique columns)
function hip_regularGridTransform(grid::_EvalRL, gridpoints::AbstractMatrix{Float64})
    r_pts = sort(unique(gridpoints[:, 1]))
    n_λ = div(size(gridpoints, 1), length(r_pts))
    return _hip_regular(grid, r_pts, gridpoints[1:n_λ, 2], nothing)
end
hip_regularGridTransform(grid::_RZGrid, gridpoints::AbstractMatrix{Float64}) =
    _hip_regular(grid, sort(unique(gridpoints[:, 1])), nothing, sort(unique(gridpoints[:, 2])))
hip_regularGridTransform(grid::_EvalRLZ, gridpoints::AbstractMatrix{Float64}) =
    _hip_regular(grid, sort(unique(gridpoints[:, 1])), sort(unique(gridpoints[:, 2])), sort(unique(gridpoints[:, 3])))

# evaluate_unstructured(source, pts; vars, out_of_bounds) (src/interpolation.jl:1735):
# the value of every variable at the rows of pts (N, ndim); out_of_bounds :nan,
# :error (DomainError) or a number; variables not in vars are NaN
function hip_evaluate_unstructured(grid::_EvalGrid, pts::AbstractMatrix{Float64}; vars = nothing,
                                   out_of_bounds = :nan)
    p = hip_plan(grid)
    gp = grid.params
    nv = length(gp.vars)
    P = Matrix{Float64}(pts)
    n = size(P, 1)
    result = Matrix{Float64}(undef, n, nv)
    fill = out_of_bounds === :nan || out_of_bounds === :error ? NaN : Float64(out_of_bounds)
    oob_error = out_of_bounds === :error ? Cint(1) : Cint(0)
    GC.@preserve P result _check(ccall((:ssg_evaluate_unstructured, libssg), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Clong, Ptr{Float64}, Cdouble, Cint, Cint, Ptr{Cvoid}),
        p.handle, pointer(grid.spectral), P, Clong(n), result, fill, oob_error, SSG_MEM_HOST, C_NULL))
    if vars !== nothing
        keep = Set(gp.vars[k] for k in intersect(vars, keys(gp.vars)))
        for v in 1:nv
            v in keep || (result[:, v] .= NaN)
        end
    end
    return result
end

# Batching form: RL / RLZ grids of one arithmetic form, device arrays fixed at
# creation, one graph launch per direction (ssg_batch_* in include/springsteel_hip.h)
mutable struct HIPBatch
    handle::Ptr{Cvoid}
    grids::Vector{Any}
end

# physical_ptrs point to Float64 arrays, or Float32 ones with precision = SSG_FP32_REFINE
function hip_batch(grids::Vector, physical_ptrs::Vector{<:Ptr}, spectral_ptrs::Vector{Ptr{Float64}};
                   precision::Cint = SSG_FP64)
    plans = Ptr{Cvoid}[hip_plan(g; precision = precision).handle for g in grids]
    phys = Ptr{Cvoid}[Ptr{Cvoid}(q) for q in physical_ptrs]
    out = Ref{Ptr{Cvoid}}(C_NULL)
    _check(ccall((:ssg_batch_create, libssg), Cint,
                 (Ptr{Ptr{Cvoid}}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Float64}}, Ref{Ptr{Cvoid}}),
                 plans, Cint(length(grids)), phys, spectral_ptrs, out))
    b = HIPBatch(out[], collect(Any, grids))
    finalizer(b) do x
        x.handle == C_NULL || ccall((:ssg_batch_destroy, libssg), Cint, (Ptr{Cvoid},), x.handle)
        x.handle = C_NULL
    end
    return b
end

hip_batch_tune!(b::HIPBatch; stream::Ptr{Cvoid} = C_NULL) =
    _check(ccall((:ssg_batch_tune, libssg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), b.handle, stream))
hip_batch_spectralTransform!(b::HIPBatch; stream::Ptr{Cvoid} = C_NULL) =
    _check(ccall((:ssg_batch_spectral_transform, libssg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), b.handle, stream))
hip_batch_gridTransform!(b::HIPBatch; stream::Ptr{Cvoid} = C_NULL) =
    _check(ccall((:ssg_batch_grid_transform, libssg), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), b.handle, stream))

# The tile pipeline of src/tiling.jl on the ring grids.  splineTransform!(shared, tile)
# (1490-1516 RL, 1675-1700 SL, 1862-1892 RLZ, 2063-2093 SLZ) writes tile.spectral, the A
# coefficients of every line, from the B coefficients in `shared` (the tile's layout);
# tileTransform!(shared, tile, physical, spectral) (1540-1658, 1720-1838, 1915-2046,
# 2113-2242) fills every physical slot from the A coefficients in `spectral` and does
# not read `shared`.  sumSpectralTile! / setSpectralTile! (1048-1167, 1204-1325) add /
# copy the tile's lines into the patch's at spectralIndexL; sumSharedSpectral's ring-grid
# method (1379-1389) copies rows 1 .. b_iDim - 3 of each variable column.  FP64 plans.
const _TileGrid = Union{_RLGrid, _RLZGrid, _SLGrid, _SLZGrid}

function hip_splineTransform!(sharedSpectral::DenseMatrix{Float64}, tile::_TileGrid; stream::Ptr{Cvoid} = C_NULL)
    _has_r3x(tile) && hip_set_ahat!(tile)
    p = hip_plan(tile)
    _check_dims(p, nothing, sharedSpectral)
    _check_dims(p, nothing, tile.spectral)
    GC.@preserve sharedSpectral tile _check(ccall((:ssg_spline_transform, libssg), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}),
        p.handle, pointer(sharedSpectral), pointer(tile.spectral), SSG_MEM_HOST, stream))
    return nothing
end

function hip_tileTransform!(sharedSpectral, tile::_TileGrid, physical::Array{Float64}, spectral::Array{Float64};
                            stream::Ptr{Cvoid} = C_NULL)
    p = hip_plan(tile)
    _check_dims(p, physical, spectral)
    GC.@preserve physical spectral _check(ccall((:ssg_tile_transform, libssg), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Cvoid}),
        p.handle, pointer(spectral), pointer(physical), SSG_MEM_HOST, stream))
    return physical
end

function _hip_tile_embed(add::Bool, patch::_TileGrid, tile::_TileGrid, stream)
    pp = hip_plan(patch)
    tp = hip_plan(tile)
    _check_dims(pp, nothing, patch.spectral)
    _check_dims(tp, nothing, tile.spectral)
    siL = Cint(tile.params.spectralIndexL)
    GC.@preserve patch tile begin
        rc = add ?
            ccall((:ssg_sum_spectral_tile, libssg), Cint,
                  (Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid}, Ptr{Float64}, Cint, Cint, Ptr{Cvoid}),
                  pp.handle, pointer(patch.spectral), tp.handle, pointer(tile.spectral), siL, SSG_MEM_HOST, stream) :
            ccall((:ssg_set_spectral_tile, libssg), Cint,
                  (Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid}, Ptr{Float64}, Cint, Cint, Ptr{Cvoid}),
                  pp.handle, pointer(patch.spectral), tp.handle, pointer(tile.spectral), siL, SSG_MEM_HOST, stream)
        _check(rc)
    end
    return patch.spectral
end

hip_sumSpectralTile!(patch::_TileGrid, tile::_TileGrid; stream::Ptr{Cvoid} = C_NULL) =
    _hip_tile_embed(true, patch, tile, stream)
hip_setSpectralTile!(patch::_TileGrid, tile::_TileGrid; stream::Ptr{Cvoid} = C_NULL) =
    _hip_tile_embed(false, patch, tile, stream)

function hip_sumSharedSpectral(sharedSpectral::DenseMatrix{Float64}, tile::_TileGrid, patchMap = nothing,
                               haloMap = nothing; stream::Ptr{Cvoid} = C_NULL)
    size(sharedSpectral, 2) == length(tile.params.vars) ||
        throw(DimensionMismatch("sharedSpectral has $(size(sharedSpectral, 2)) columns, the tile $(length(tile.params.vars)) variables"))
    p = hip_plan(tile)
    GC.@preserve sharedSpectral tile _check(ccall((:ssg_sum_shared_spectral, libssg), Cint,
        (Ptr{Float64}, Clong, Ptr{Cvoid}, Ptr{Float64}, Cint, Cint, Ptr{Cvoid}),
        pointer(sharedSpectral), Clong(size(sharedSpectral, 1)), p.handle, pointer(tile.spectral),
        Cint(tile.params.spectralIndexL), SSG_MEM_HOST, stream))
    return nothing
end

export hip_spectralTransform!, hip_gridTransform!, hip_spectralTransform, hip_gridTransform,
       hip_spectralTransform_device, hip_gridTransform_device, hip_plan, hip_tune!, hip_tuning, hip_set_tuning!,
       hip_batch, hip_batch_tune!, hip_batch_spectralTransform!, hip_batch_gridTransform!, hip_regularGridTransform,
       hip_evaluate_unstructured, hip_set_ahat!, hip_applyFilter!, hip_splineTransform!, hip_tileTransform!,
       hip_sumSpectralTile!, hip_setSpectralTile!, hip_sumSharedSpectral

end # module
