# QInchwormAMD.jl -- ccall binding of libqiw (include/qiw.h) for QInchworm.jl.
#
# Drop-in bodies for `inchworm!` and `correlator_2p` (src/inchworm.jl:332-498, :1075-1084):
# the caller keeps building `Expansion` (KeldyshED/Keldysh) and calls these functions with
# the reference's signatures; sampling and diagram evaluation run on one MI355X.
#
# Conventions: sector indices are passed 0-based; complex arrays are ComplexF64 (interleaved,
# identical to C99 double _Complex); block matrices column-major; P tables in "BT layout"
# (per grid point, all sector blocks concatenated).
module QInchwormAMD

using Keldysh; kd = Keldysh
using KeldyshED
using QInchworm.expansion: Expansion
using QInchworm.spline_gf: SplineInterpolatedGF, IncSplineImaginaryTimeGF, make_inc_interpolant
using QInchworm.utility: extend!
using QInchworm.ppgf: set_ppgf!
using QInchworm.sector_block_matrix: SectorBlockMatrix
using QInchworm.scrambled_sobol: ScrambledSobolSeq, next!
using QInchworm.randomization: RandomizationParams, RequestStdDev
using QInchworm.keldysh_dlr: DLRImaginaryTimeGF

const libqiw = get(ENV, "QIW_LIBRARY", joinpath(@__DIR__, "..", "lib", "libqiw.so"))

const QIW_PREC_FP64 = Int32(0)
const QIW_PREC_FP32 = Int32(1)
const QIW_GF_GRID_LINEAR = Int32(0)
const QIW_GF_CUBIC_SPLINE = Int32(1)
const QIW_GF_INC_SPLINE = Int32(2)
const QIW_GF_DLR = Int32(3)

struct QiwSbm
    n_blocks::Int32
    s_i::Ptr{Int32}
    s_f::Ptr{Int32}
    data::Ptr{ComplexF64}
end

struct QiwModel
    beta::Float64
    n_tau::Int32
    n_sectors::Int32
    sector_dim::Ptr{Int32}
    energies::Ptr{Float64}
    ppgf_kind::Int32
    n_pairs::Int32
    pair_op_i::Ptr{QiwSbm}
    pair_op_f::Ptr{QiwSbm}
    pair_prop::Ptr{ComplexF64}
    pair_prop_kind::Ptr{Int32}
    pair_prop_xi::Ptr{Int32}
    n_corr::Int32
    corr_A::Ptr{QiwSbm}
    corr_B::Ptr{QiwSbm}
    dlr_rank::Int32                 # QIW_GF_DLR pairs: frequencies per propagator
    pair_dlr_omega::Ptr{Float64}    # [n_pairs][dlr_rank]
    pair_dlr_coef::Ptr{ComplexF64}  # [n_pairs][dlr_rank]
end

# Randomized qMC (qiw_rand).  With rand_params.rng given and target_std = 0 the sequences
# are the reference's own (scramble = 2): stage_scrambles! constructs the
# ScrambledSobolSeq objects the reference's loops would, in their order, from rng.  With a
# stopping target the scramble bits come from libqiw's counter-based generator keyed by a
# seed drawn from rng once per call (scramble = 1).  Any other seq_type (the reference's
# sequence plug point, src/randomization.jl:82-94; e.g. RandomSeq, src/utility.jl:336-414)
# is constructed the same way and its points are staged (scramble = 3).
struct QiwRand
    N_seqs::Int32
    scramble::Int32
    seed::UInt64
    target_std::Float64
end

caller_drawn(rp::RandomizationParams) = rp.rng !== nothing && rp.target_std == 0.0

function QiwRand(rp::RandomizationParams, seq_type::Type = ScrambledSobolSeq)
    if seq_type !== ScrambledSobolSeq
        rp.target_std == 0.0 || throw(ArgumentError("a seq_type other than ScrambledSobolSeq needs target_std = 0"))
        return QiwRand(Int32(rp.N_seqs), Int32(3), UInt64(0), 0.0)
    end
    return QiwRand(Int32(rp.N_seqs), rp.rng === nothing ? Int32(0) : caller_drawn(rp) ? Int32(2) : Int32(1),
                   (rp.rng === nothing || caller_drawn(rp)) ? UInt64(0) : rand(rp.rng, UInt64), rp.target_std)
end

const QIW_CALL_INCHWORM = Int32(0)
const QIW_CALL_CORRELATOR = Int32(3)
const QIW_CALL_DIFF = Int32(4)

struct QiwRun
    n_orders::Int32
    orders::Ptr{Int32}
    n_orders_bare::Int32
    orders_bare::Ptr{Int32}
    N_samples::Int64
    n_pts_after_max::Int32
    precision::Int32
    rand_params::Ptr{QiwRand}   # C_NULL: RandomizationParams()
end

# The sequences of one call under scramble = 2: qiw_scramble_dims lists their dimensions in
# the order the reference constructs them (src/randomization.jl:94); each is drawn from rng
# exactly as the reference draws it and staged as (m[i][d][j], x[i][d]).
# Under scramble = 3 the sequences are seq_type(D; scramble_rng = rng), constructed in the
# same order, and their first N_samples points (next!, src/qmc_integrate.jl:597-612) are
# staged as x[d, r, i].
function stage_scrambles!(ctx, run::Ref{QiwRun}, call::Int32, rp::RandomizationParams,
                          seq_type::Type = ScrambledSobolSeq)
    points = seq_type !== ScrambledSobolSeq
    (points || caller_drawn(rp)) || return nothing
    sig = (Ptr{Cvoid}, Ref{QiwRun}, Int32, Int64, Ptr{Int32})
    n = ccall((:qiw_scramble_dims, libqiw), Int64, sig, ctx.h, run, call, 0, C_NULL)
    n < 0 && check(ctx, Cint(n))
    dims = Vector{Int32}(undef, n)
    check(ctx, Cint(min(0, ccall((:qiw_scramble_dims, libqiw), Int64, sig, ctx.h, run, call, n, dims))))
    stride = Int(maximum(dims; init = Int32(2)))
    if points
        N = max(Int(run[].N_samples), 1)
        x = zeros(Float64, stride, N, n)
        for (i, D) in enumerate(dims)
            seq = seq_type(Int(D); scramble_rng = rp.rng)
            for r in 1:N
                x[1:D, r, i] = next!(seq)
            end
        end
        check(ctx, ccall((:qiw_set_points, libqiw), Cint, (Ptr{Cvoid}, Int64, Int64, Int32, Ptr{Float64}),
                         ctx.h, n, N, Int32(stride), x))
        return nothing
    end
    m = zeros(UInt32, 32, stride, n)
    x = zeros(UInt32, stride, n)
    for (i, D) in enumerate(dims)
        seq = ScrambledSobolSeq(Int(D); scramble_rng = rp.rng)
        m[:, 1:D, i] = permutedims(seq.m)
        x[1:D, i] = seq.x
    end
    check(ctx, ccall((:qiw_set_scrambles, libqiw), Cint, (Ptr{Cvoid}, Int64, Int32, Ptr{UInt32}, Ptr{UInt32}),
                     ctx.h, n, Int32(stride), m, x))
    return nothing
end

# Learned warp of the unit cube (qiw_warp_config); no counterpart in the reference
struct QiwWarpConfig
    n_layers::Int32
    n_bins::Int32
    hidden::Int32
    iters::Int32
    batch::Int32
    n_train_steps::Int32
    n_train::Int64
    lr::Float64
    seed::UInt64
end
QiwWarpConfig(; n_layers = 4, n_bins = 8, hidden = 64, iters = 4000, batch = 8192, n_train_steps = 8,
              n_train = 2^14, lr = 2e-3, seed = 1) =
    QiwWarpConfig(n_layers, n_bins, hidden, iters, batch, n_train_steps, n_train, lr, seed)

# Work mapping of the evaluators (qiw_tuning); no counterpart in the reference
struct QiwTuning
    chunk_words::Int32
    samples_per_task::Int32
    chunk_cost::Int64
    n_tried::Int32
    pad_::Int32
    best_ms::Float64
end
QiwTuning(; chunk_words = 0, samples_per_task = 0, chunk_cost = 0) =
    QiwTuning(chunk_words, samples_per_task, chunk_cost, 0, 0, 0.0)

mutable struct Context
    h::Ptr{Cvoid}
end

function check(ctx::Context, rc)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:qiw_last_error, libqiw), Cstring, (Ptr{Cvoid},), ctx.h))
    error("libqiw error $(rc): $(msg)")
end

function Context(device::Integer = 0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:qiw_create, libqiw), Cint, (Ref{Ptr{Cvoid}}, Cint), h, device)
    rc == 0 || error("qiw_create failed with code $(rc)")
    ctx = Context(h[])
    finalizer(c -> ccall((:qiw_destroy, libqiw), Cvoid, (Ptr{Cvoid},), c.h), ctx)
    return ctx
end

const default_ctx = Ref{Union{Context, Nothing}}(nothing)
context() = something(default_ctx[], (default_ctx[] = Context()))

# SectorBlockMatrix -> (s_i, s_f, data) arrays kept alive by `keep`
function sbm_arrays(sbm, keep)
    keys_sorted = sort(collect(keys(sbm)))
    s_i = Int32[k - 1 for k in keys_sorted]
    s_f = Int32[sbm[k][1] - 1 for k in keys_sorted]
    data = isempty(keys_sorted) ? ComplexF64[0] : reduce(vcat, [vec(sbm[k][2]) for k in keys_sorted])
    push!(keep, s_i, s_f, data)
    return QiwSbm(length(s_i), pointer(s_i), pointer(s_f), pointer(data))
end

# Grid values and interpolation kind of an InteractionPair propagator
function prop_table(g, n_tau)
    if g isa SplineInterpolatedGF
        return vec(g.GF.mat.data)[1:n_tau], QIW_GF_CUBIC_SPLINE, Int32(Int(g.GF.ξ))
    end
    if g isa DLRImaginaryTimeGF  # pole sum; the grid values are not read (zeros)
        return zeros(ComplexF64, n_tau), QIW_GF_DLR, Int32(Int(g.ξ))
    end
    return vec(g.mat.data)[1:n_tau], QIW_GF_GRID_LINEAR, Int32(Int(g.ξ))
end

# (frequencies, coefficients) of the DLR pair propagators, [n_pairs][rank] row-major
# (zero rows for the other pairs): the DLR grid's frequencies G.grid.dlr.ω and the
# coefficients G.mat.data that Lehmann.jl's dlr2tau sums (src/keldysh_dlr.jl:172-176).
function dlr_tables(pairs)
    dl = [p.propagator for p in pairs if p.propagator isa DLRImaginaryTimeGF]
    isempty(dl) && return 0, Float64[], ComplexF64[]
    rank = maximum(length(g.grid.dlr.ω) for g in dl)
    w = zeros(Float64, rank, length(pairs))
    c = zeros(ComplexF64, rank, length(pairs))
    for (i, p) in enumerate(pairs)
        g = p.propagator
        g isa DLRImaginaryTimeGF || continue
        k = length(g.grid.dlr.ω)
        w[1:k, i] = g.grid.dlr.ω
        c[1:k, i] = vec(g.mat.data)
    end
    return rank, vec(w), vec(c)
end

function stage!(ctx::Context, exp::Expansion, grid::kd.ImaginaryTimeGrid, max_order::Int, keep)
    n_tau = length(grid)
    dims = Int32[length(s) for s in exp.ed.subspaces]
    energies = reduce(vcat, [Float64.(e) for e in KeldyshED.energies(exp.ed)])
    ppgf_kind = first(exp.P) isa IncSplineImaginaryTimeGF ? QIW_GF_INC_SPLINE : QIW_GF_GRID_LINEAR
    op_i = [sbm_arrays(p[1], keep) for p in exp.pair_operator_mat]
    op_f = [sbm_arrays(p[2], keep) for p in exp.pair_operator_mat]
    tabs = [prop_table(p.propagator, n_tau) for p in exp.pairs]
    props = reduce(vcat, [t[1] for t in tabs])
    kinds = Int32[t[2] for t in tabs]
    xis = Int32[t[3] for t in tabs]
    cA = [sbm_arrays(p[1], keep) for p in exp.corr_operators_mat]
    cB = [sbm_arrays(p[2], keep) for p in exp.corr_operators_mat]
    rank, dw, dc = dlr_tables(exp.pairs)
    push!(keep, dims, energies, op_i, op_f, props, kinds, xis, cA, cB, dw, dc)
    model = Ref(QiwModel(grid.contour.β, n_tau, length(dims), pointer(dims), pointer(energies), ppgf_kind,
                         length(exp.pairs), pointer(op_i), pointer(op_f), pointer(props), pointer(kinds),
                         pointer(xis), length(cA), isempty(cA) ? C_NULL : pointer(cA),
                         isempty(cB) ? C_NULL : pointer(cB), Int32(rank),
                         rank == 0 ? C_NULL : pointer(dw), rank == 0 ? C_NULL : pointer(dc)))
    check(ctx, ccall((:qiw_set_model, libqiw), Cint, (Ptr{Cvoid}, Ref{QiwModel}), ctx.h, model))
    # Sobol state of the reference's own generator (unscrambled by default)
    D = max(2 * max_order, 1)
    seq = ScrambledSobolSeq(D)
    m = Matrix{UInt32}(transpose(seq.m))          # m[d*32 + j] = seq.m[d+1, j+1]
    x = copy(seq.x)
    push!(keep, m, x)
    check(ctx, ccall((:qiw_set_sobol, libqiw), Cint, (Ptr{Cvoid}, Int32, Ptr{UInt32}, Ptr{UInt32}),
                     ctx.h, D, m, x))
end

# BT-layout table of expansion.P on the grid
function ppgf_table(exp::Expansion, n_tau)
    bt = sum(length(s)^2 for s in exp.ed.subspaces)
    T = zeros(ComplexF64, bt, n_tau)
    for j in 1:n_tau
        off = 0
        for P_s in exp.P
            GF = P_s isa IncSplineImaginaryTimeGF ? P_s.GF : P_s
            blk = GF.mat.data[:, :, j]
            T[off+1:off+length(blk), j] = vec(blk)
            off += length(blk)
        end
    end
    return T
end

# expansion.P after a run, read back from the context: the grid values of every sector and,
# for IncSplineImaginaryTimeGF, the incremental interpolants as the reference leaves them
# (normalize! rebuilds them from the derivative at 0 over the known rows,
# src/ppgf.jl:677-695).
function load_ppgf!(ctx::Context, exp::Expansion, grid)
    n_tau = length(grid)
    bt = sum(length(s)^2 for s in exp.ed.subspaces)
    T = zeros(ComplexF64, bt, n_tau)
    der0 = zeros(ComplexF64, bt)
    n_known = Ref{Int32}(0)
    check(ctx, ccall((:qiw_get_ppgf, libqiw), Cint, (Ptr{Cvoid}, Ref{Int32}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                     ctx.h, n_known, T, der0))
    off = 0
    for P_s in exp.P
        GF = P_s isa IncSplineImaginaryTimeGF ? P_s.GF : P_s
        n = size(GF.mat.data, 1)
        for (j, τ) in enumerate(grid)
            GF[τ, grid[1]] = reshape(T[off+1:off+n*n, j], n, n)
        end
        if P_s isa IncSplineImaginaryTimeGF
            d0 = reshape(der0[off+1:off+n*n], n, n)
            for k in 1:n, l in 1:n
                P_s.interpolants[k, l] = make_inc_interpolant(GF, k, l, d0[k, l])
                for j in 2:n_known[]
                    extend!(P_s.interpolants[k, l], GF.mat.data[k, l, j])
                end
            end
        end
        off += n * n
    end
end

# Row j of a BT-layout table as a SectorBlockMatrix (block-diagonal, sector s -> s)
function sbm_row(exp::Expansion, T, j)::SectorBlockMatrix
    out = SectorBlockMatrix()
    off = 0
    for (s, sub) in enumerate(exp.ed.subspaces)
        n = length(sub)
        out[s] = (s, reshape(T[off+1:off+n*n, j], n, n))
        off += n * n
    end
    return out
end

"""
    inchworm!(expansion, grid, orders, orders_bare, N_samples; n_pts_after_max, rand_params)

Drop-in for QInchworm.inchworm! on one MI355X.  Writes expansion.P and returns
(P_orders, P_orders_std) as the reference does (src/inchworm.jl:369-370, :413-416,
:487-491): Dict(order => kd.zero(expansion.P)) filled by set_ppgf! with the order's
contribution of the initial step (orders_bare) and of every regular step (orders), in
step order.  With rand_params.N_seqs > 1 every group is integrated with up to N_seqs
scrambled sequences (with target_std = 0 the reference's own, drawn from rand_params.rng
in its order) and P_orders_std holds the standard deviations.
"""
function inchworm!(exp::Expansion, grid::kd.ImaginaryTimeGrid, orders, orders_bare, N_samples::Int64;
                   n_pts_after_max::Int64 = typemax(Int64), rand_params::RandomizationParams = RandomizationParams(),
                   seq_type::Type = ScrambledSobolSeq, precision::Int32 = QIW_PREC_FP64, ctx::Context = context())
    @assert N_samples == 0 || ispow2(N_samples)
    @assert rand_params.N_seqs > 0
    rq = Ref(QiwRand(rand_params, seq_type))
    keep = Any[]
    orders_v = Int32.(collect(orders)); bare_v = Int32.(collect(orders_bare))
    stage!(ctx, exp, grid, Int(maximum(vcat(orders_v, bare_v, Int32[0]))), keep)
    n_tau = length(grid)
    bt = sum(length(s)^2 for s in exp.ed.subspaces)
    oall = sort(unique(vcat(orders_v, bare_v)))
    P = zeros(ComplexF64, bt, n_tau)
    Po = zeros(ComplexF64, bt, n_tau, length(oall))
    Ps = zeros(ComplexF64, bt, n_tau, length(oall))
    GC.@preserve keep orders_v bare_v rq begin
        run = Ref(QiwRun(length(orders_v), pointer(orders_v), length(bare_v), pointer(bare_v), N_samples,
                         n_pts_after_max == typemax(Int64) ? 0 : Int32(n_pts_after_max), precision,
                         Base.unsafe_convert(Ptr{QiwRand}, rq)))
        stage_scrambles!(ctx, run, QIW_CALL_INCHWORM, rand_params, seq_type)
        check(ctx, ccall((:qiw_inchworm, libqiw), Cint,
                         (Ptr{Cvoid}, Ref{QiwRun}, Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                         ctx.h, run, P, Po, Ps))
    end
    load_ppgf!(ctx, exp, grid)
    P_orders = Dict(Int(o) => kd.zero(exp.P) for o in oall)
    P_orders_std = Dict(Int(o) => kd.zero(exp.P) for o in oall)
    slot = Dict(Int(o) => i for (i, o) in enumerate(oall))
    for o in unique(bare_v)          # initial step: grid[1] -> grid[2]
        set_ppgf!(P_orders[o], grid[1], grid[2], sbm_row(exp, view(Po, :, :, slot[o]), 2))
        set_ppgf!(P_orders_std[o], grid[1], grid[2], sbm_row(exp, view(Ps, :, :, slot[o]), 2))
    end
    for n in 2:n_tau-1, o in unique(orders_v)   # regular steps: grid[1] -> grid[n + 1]
        set_ppgf!(P_orders[o], grid[1], grid[n + 1], sbm_row(exp, view(Po, :, :, slot[o]), n + 1))
        set_ppgf!(P_orders_std[o], grid[1], grid[n + 1], sbm_row(exp, view(Ps, :, :, slot[o]), n + 1))
    end
    return P_orders, P_orders_std
end

"""
    diff_inchworm!(expansion, grid, orders, N_samples; rand_params)

Drop-in for QInchworm.diff_inchworm! (src/inchworm.jl:666-778) on one MI355X.  Writes
expansion.P and returns the order-resolved self-energy and its standard deviations as
Dict{Int, Matrix{ComplexF64}} in BT layout (sum m^2 x n_tau); the reference returns
nothing, so a caller of the reference's signature ignores them.
"""
function diff_inchworm!(exp::Expansion, grid::kd.ImaginaryTimeGrid, orders, N_samples::Int64;
                        rand_params::RandomizationParams = RandomizationParams(),
                        seq_type::Type = ScrambledSobolSeq, precision::Int32 = QIW_PREC_FP64,
                        ctx::Context = context())
    @assert N_samples == 0 || ispow2(N_samples)
    @assert rand_params.N_seqs > 0
    rq = Ref(QiwRand(rand_params, seq_type))
    keep = Any[]
    orders_v = Int32.(collect(orders))
    stage!(ctx, exp, grid, Int(maximum(vcat(orders_v, Int32[0]))), keep)
    n_tau = length(grid)
    bt = sum(length(s)^2 for s in exp.ed.subspaces)
    P = zeros(ComplexF64, bt, n_tau)
    So = zeros(ComplexF64, bt, n_tau, length(orders_v))
    Ss = zeros(ComplexF64, bt, n_tau, length(orders_v))
    GC.@preserve keep orders_v rq begin
        run = Ref(QiwRun(length(orders_v), pointer(orders_v), 0, C_NULL, N_samples, 0, precision,
                         Base.unsafe_convert(Ptr{QiwRand}, rq)))
        stage_scrambles!(ctx, run, QIW_CALL_DIFF, rand_params, seq_type)
        check(ctx, ccall((:qiw_diff_inchworm, libqiw), Cint,
                         (Ptr{Cvoid}, Ref{QiwRun}, Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{ComplexF64}),
                         ctx.h, run, P, So, Ss))
    end
    load_ppgf!(ctx, exp, grid)
    # Σ_orders as the reference builds them (src/inchworm.jl:698-699, :770-771): containers
    # kd.zero(expansion.P) filled with set_ppgf! in step order.  The rows are where the
    # reference's set_ppgf! leaves each step's value: every row of a grid container, rows
    # 2..n_tau (appended one per step) of an incremental one.
    Σ_orders = Dict(Int(o) => kd.zero(exp.P) for o in orders_v)
    Σ_orders_std = Dict(Int(o) => kd.zero(exp.P) for o in orders_v)
    first_row = first(exp.P) isa IncSplineImaginaryTimeGF ? 2 : 1
    for n in first_row:n_tau, (i, o) in enumerate(orders_v)
        set_ppgf!(Σ_orders[Int(o)], grid[1], grid[n], sbm_row(exp, view(So, :, :, i), n))
        set_ppgf!(Σ_orders_std[Int(o)], grid[1], grid[n], sbm_row(exp, view(Ss, :, :, i), n))
    end
    return Σ_orders, Σ_orders_std
end

"""
    correlator_2p(expansion, grid, orders, N_samples; rand_params)
    correlator_2p(expansion, grid, orders, N_samples, RequestStdDev(); rand_params)

Drop-ins for QInchworm.correlator_2p on one MI355X (src/inchworm.jl:1075-1084 and
:918-1052): a Vector{kd.ImaginaryTimeGF{ComplexF64, true}} with <A(tau) B(0)> for every
pair in expansion.corr_operators, fermionic when A and B both are (:996-1004); the
RequestStdDev method returns the tuple (values, standard deviations).
"""
function correlator_2p(exp::Expansion, grid::kd.ImaginaryTimeGrid, orders, N_samples::Int64;
                       rand_params::RandomizationParams = RandomizationParams(),
                       seq_type::Type = ScrambledSobolSeq,
                       ctx::Context = context())::Vector{kd.ImaginaryTimeGF{ComplexF64, true}}
    return correlator_2p_impl(exp, grid, orders, N_samples, rand_params, seq_type, false, ctx)[1]
end

function correlator_2p(exp::Expansion, grid::kd.ImaginaryTimeGrid, orders, N_samples::Int64, ::RequestStdDev;
                       rand_params::RandomizationParams = RandomizationParams(),
                       seq_type::Type = ScrambledSobolSeq, ctx::Context = context())
    return correlator_2p_impl(exp, grid, orders, N_samples, rand_params, seq_type, true, ctx)
end

function correlator_2p_impl(exp::Expansion, grid::kd.ImaginaryTimeGrid, orders, N_samples::Int64,
                            rand_params::RandomizationParams, seq_type::Type, with_std::Bool, ctx::Context)
    @assert N_samples == 0 || ispow2(N_samples)
    @assert rand_params.N_seqs > 0
    rq = Ref(QiwRand(rand_params, seq_type))
    keep = Any[]
    orders_v = Int32.(collect(orders))
    stage!(ctx, exp, grid, Int(maximum(vcat(orders_v, Int32[0]))), keep)
    n_tau = length(grid)
    T = ppgf_table(exp, n_tau)
    der0 = zeros(ComplexF64, size(T, 1))
    check(ctx, ccall((:qiw_set_ppgf, libqiw), Cint, (Ptr{Cvoid}, Int32, Ptr{ComplexF64}, Ptr{ComplexF64}),
                     ctx.h, n_tau, T, der0))
    out = zeros(ComplexF64, n_tau, length(exp.corr_operators))
    out_std = zeros(ComplexF64, n_tau, length(exp.corr_operators))
    GC.@preserve keep orders_v rq begin
        run = Ref(QiwRun(length(orders_v), pointer(orders_v), 0, C_NULL, N_samples, 0, 0,
                         Base.unsafe_convert(Ptr{QiwRand}, rq)))
        stage_scrambles!(ctx, run, QIW_CALL_CORRELATOR, rand_params, seq_type)
        check(ctx, ccall((:qiw_correlator_2p_std, libqiw), Cint,
                         (Ptr{Cvoid}, Ref{QiwRun}, Ptr{ComplexF64}, Ptr{ComplexF64}), ctx.h, run, out, out_std))
    end
    corr_list = kd.ImaginaryTimeGF{ComplexF64, true}[]
    corr_std_list = kd.ImaginaryTimeGF{ComplexF64, true}[]
    is_fermion(expr) = isodd(length(first(keys(expr.monomials))))
    for (c, (A, B)) in enumerate(exp.corr_operators)
        @assert length(A) == 1 "Operator A must be a single monomial in C/C^+"
        @assert length(B) == 1 "Operator B must be a single monomial in C/C^+"
        ξ = (is_fermion(A) && is_fermion(B)) ? kd.fermionic : kd.bosonic
        g = kd.ImaginaryTimeGF(grid, 1, ξ, true)
        g_std = kd.ImaginaryTimeGF(grid, 1, ξ, true)
        for (n, τ) in enumerate(grid)
            g[τ, grid[1]] = out[n, c]
            g_std[τ, grid[1]] = out_std[n, c]
        end
        push!(corr_list, g)
        push!(corr_std_list, g_std)
    end
    return (corr_list, corr_std_list)
end

"""
    train_warp!(expansion, grid, orders; config = QiwWarpConfig(), ctx)

Fit the learned warp to |integrand| samples of the plain method at `expansion.P` (run a
plain `inchworm!` first) and enable it for the following `inchworm!` calls on `ctx`.
Returns the loss of every Adam iteration.
"""
function train_warp!(exp::Expansion, grid::kd.ImaginaryTimeGrid, orders; config::QiwWarpConfig = QiwWarpConfig(),
                     ctx::Context = context())
    keep = Any[]
    orders_v = Int32.(collect(orders))
    stage!(ctx, exp, grid, Int(maximum(vcat(orders_v, Int32[0]))), keep)
    n_tau = length(grid)
    T = ppgf_table(exp, n_tau)
    der0 = zeros(ComplexF64, size(T, 1))
    check(ctx, ccall((:qiw_set_ppgf, libqiw), Cint, (Ptr{Cvoid}, Int32, Ptr{ComplexF64}, Ptr{ComplexF64}),
                     ctx.h, n_tau, T, der0))
    cfg = Ref(config)
    check(ctx, ccall((:qiw_warp_init, libqiw), Cint, (Ptr{Cvoid}, Ref{QiwWarpConfig}), ctx.h, cfg))
    loss = zeros(Float64, max(config.iters, 1))
    run = Ref(QiwRun(length(orders_v), pointer(orders_v), 0, C_NULL, 0, 0, QIW_PREC_FP64, C_NULL))
    GC.@preserve keep orders_v begin
        check(ctx, ccall((:qiw_warp_train, libqiw), Cint, (Ptr{Cvoid}, Ref{QiwRun}, Ptr{Float64}),
                         ctx.h, run, loss))
    end
    check(ctx, ccall((:qiw_warp_enable, libqiw), Cint, (Ptr{Cvoid}, Int32), ctx.h, 1))
    return loss[1:config.iters]
end

"""
    disable_warp!(ctx)

Return to plain quasi Monte Carlo points.
"""
disable_warp!(ctx::Context = context()) =
    check(ctx, ccall((:qiw_warp_enable, libqiw), Cint, (Ptr{Cvoid}, Int32), ctx.h, 0))

"""
    set_tuning!(tuning = QiwTuning(); ctx)

Fix the work mapping of the evaluators (topology chunk size, samples per workgroup;
0 keeps a default).  Results agree with any setting to rounding.
"""
set_tuning!(t::QiwTuning = QiwTuning(); ctx::Context = context()) =
    check(ctx, ccall((:qiw_set_tuning, libqiw), Cint, (Ptr{Cvoid}, Ref{QiwTuning}), ctx.h, Ref(t)))

"""
    set_graphs!(on::Bool; ctx)

Replay the bold-step loop of `inchworm!` from a hipGraph (the default; results are
identical) or launch every kernel from the host (`qiw_set_graphs`).
"""
set_graphs!(on::Bool; ctx::Context = context()) =
    check(ctx, ccall((:qiw_set_graphs, libqiw), Cint, (Ptr{Cvoid}, Int32), ctx.h, Int32(on)))

"""
    autotune!(expansion, grid, orders, N_samples; precision = QIW_PREC_FP64, ctx)

Time candidate work mappings on one bold step at the expansion's current `P` (tau_w at
the middle of the grid) and keep the fastest for the following calls on `ctx`.  Returns
the kept `QiwTuning` (with the number of candidates and the best launch time).
"""
function autotune!(exp::Expansion, grid::kd.ImaginaryTimeGrid, orders, N_samples::Integer;
                   precision::Int32 = QIW_PREC_FP64, ctx::Context = context())
    keep = Any[]
    orders_v = Int32.(collect(orders))
    stage!(ctx, exp, grid, Int(maximum(vcat(orders_v, Int32[0]))), keep)
    n_tau = length(grid)
    T = ppgf_table(exp, n_tau)
    der0 = zeros(ComplexF64, size(T, 1))
    check(ctx, ccall((:qiw_set_ppgf, libqiw), Cint, (Ptr{Cvoid}, Int32, Ptr{ComplexF64}, Ptr{ComplexF64}),
                     ctx.h, n_tau, T, der0))
    out = Ref(QiwTuning())
    run = Ref(QiwRun(length(orders_v), pointer(orders_v), 0, C_NULL, N_samples, 0, precision, C_NULL))
    GC.@preserve keep orders_v begin
        check(ctx, ccall((:qiw_autotune, libqiw), Cint, (Ptr{Cvoid}, Ref{QiwRun}, Ref{QiwTuning}),
                         ctx.h, run, out))
    end
    return out[]
end

end # module QInchwormAMD
