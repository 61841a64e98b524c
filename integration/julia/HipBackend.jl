# HipBackend.jl — HaploADMIXTURE.jl GPU plugin over libhaplo_hip.so (MI355X / gfx950).
#
# Included into module HaploADMIXTURE in place of src/cuda/{structs,kernels,runners}.jl (see the edit to
# src/HaploADMIXTURE.jl:13-21 in INTEGRATION.md).  It defines the names the reference's host code calls when
# d_cu !== nothing, so that src/algorithms_outer.jl and src/algorithms_inner.jl run unchanged:
#
#   _cu_admixture_base(d, g_la, I, J) -> (d_cu, g_cu)           driver.jl:153-158 (replaces cuda/structs.jl:34-44)
#   d_cu.q, d_cu.p, d_cu.q_next, d_cu.p_next,                   cuda/structs.jl:1-15: device arrays, here HipBuf
#   d_cu.XtX_q, d_cu.Xtz_q, d_cu.XtX_p, d_cu.Xtz_p              proxies of the context's named arrays
#   OpenADMIXTURE.copyto_sync!([d_cu.q, d_cu.p], [q, p])        algorithms_inner.jl:36,80,124,182,250,326;
#   OpenADMIXTURE.copyto_sync!([q, p], [d_cu.q_next, ...])       algorithms_inner.jl:42,126,252; outer :50,122,149
#   d_cu.p .= d_cu.p_next                                        algorithms_outer.jl:9-10
#   loglikelihood(d_cu, g_cu, d_cu.q, d_cu.p)                   algorithms_outer.jl:11,51,123,150; inner :183,327
#   em!(d_cu, g_cu), em_p!(d_cu, g_cu)                          algorithms_inner.jl:38,40 (cuda/runners.jl:89-99)
#   update_q_cuda!(d_cu, g_cu), update_p_cuda!(d_cu, g_cu)      algorithms_inner.jl:125,251 (runners.jl:115-136)
#
# g_cu is `nothing`: the genotypes live inside the library's context (packed once at _cu_admixture_base).
# The QP, projections and ZAL stay in Julia (OpenADMIXTURE), exactly as with the CUDA plugin.  hip_fit!
# instead runs init_em! + admixture_qn! entirely on the device (hk_fit).
#
# Every call blocks until its results are on the host (the library synchronises its stream), as CUDA.@sync
# does in cuda/runners.jl.  Arrays cross in FP64; host arrays of another element type (T2 = Float32) are
# converted on the way.

using Libdl

const libhk = get(ENV, "HAPLO_HIP_LIB", "libhaplo_hip")

# hk_buffer (include/haplo_hip.h)
const HK_BUF_Q, HK_BUF_P, HK_BUF_Q_NEXT, HK_BUF_P_NEXT = Cint(0), Cint(1), Cint(2), Cint(3)
const HK_BUF_XTX_Q, HK_BUF_XTZ_Q, HK_BUF_XTX_P, HK_BUF_XTZ_P = Cint(4), Cint(5), Cint(6), Cint(7)
const HK_PREC_FP64, HK_PREC_MIXED, HK_PREC_FP32 = Cint(0), Cint(1), Cint(2)

hk_error(h::Ptr{Cvoid}) = unsafe_string(ccall((:hk_last_error, libhk), Cstring, (Ptr{Cvoid},), h))
hk_check(h::Ptr{Cvoid}, rc::Cint) = rc == 0 || error("libhaplo_hip (status $rc): ", hk_error(h))

# ---- device arrays ---------------------------------------------------------------------------------------

"""
    HipBuf{N}

A named device array of a libhaplo_hip context (q, p, q_next, p_next, XtX_q, Xtz_q, XtX_p, Xtz_p), standing in
for the CuArray fields of CuAdmixData.  It has a size but no host-side elements: data moves only through
`copyto_sync!` and `dst .= src` between device arrays.
"""
struct HipBuf{N} <: AbstractArray{Float64, N}
    h::Ptr{Cvoid}
    id::Cint
    dims::NTuple{N, Int}
end
Base.size(b::HipBuf) = b.dims
Base.IndexStyle(::Type{<:HipBuf}) = IndexLinear()
Base.getindex(::HipBuf, ::Int) = error("HipBuf is device memory: copy it to the host with copyto_sync!")
Base.setindex!(::HipBuf, _, ::Int) = error("HipBuf is device memory: copy to it with copyto_sync!")
Base.show(io::IO, b::HipBuf) = print(io, "HipBuf(", join(b.dims, "×"), ", id ", b.id, ")")
Base.show(io::IO, ::MIME"text/plain", b::HipBuf) = show(io, b)

# d_cu.p .= d_cu.p_next: device-to-device copy (hk_buf_copy)
struct HipStyle <: Broadcast.AbstractArrayStyle{Any} end
HipStyle(::Val) = HipStyle()
Base.BroadcastStyle(::Type{<:HipBuf}) = HipStyle()
function Base.copyto!(dst::HipBuf, bc::Broadcast.Broadcasted{HipStyle})
    args = bc.args
    (bc.f === identity && length(args) == 1 && args[1] isa HipBuf) ||
        error("HipBuf supports only `dst .= src` between device arrays of one context")
    src = args[1]
    src.h == dst.h || error("HipBuf copy across contexts")
    hk_check(dst.h, ccall((:hk_buf_copy, libhk), Cint, (Ptr{Cvoid}, Cint, Cint), dst.h, dst.id, src.id))
    dst
end

function _upload!(dst::HipBuf, src::AbstractArray)
    size(src) == size(dst) || throw(DimensionMismatch("copy of $(size(src)) into $(size(dst))"))
    a = src isa Array{Float64} ? src : Array{Float64}(src)
    hk_check(dst.h, ccall((:hk_buf_upload, libhk), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}), dst.h, dst.id, a))
end

function _download!(dst::AbstractArray, src::HipBuf)
    size(src) == size(dst) || throw(DimensionMismatch("copy of $(size(src)) into $(size(dst))"))
    if dst isa Array{Float64}
        hk_check(src.h, ccall((:hk_buf_download, libhk), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}), src.h, src.id, dst))
    else
        tmp = Array{Float64}(undef, size(src))
        hk_check(src.h, ccall((:hk_buf_download, libhk), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}), src.h, src.id, tmp))
        dst .= tmp
    end
end

# OpenADMIXTURE.copyto_sync!(dsts, srcs): element-wise copies between host arrays and device arrays, in order.
function OpenADMIXTURE.copyto_sync!(dsts::Vector{<:HipBuf}, srcs::Vector{<:AbstractArray})
    length(dsts) == length(srcs) || throw(ArgumentError("copyto_sync!: lengths differ"))
    for (d, s) in zip(dsts, srcs)
        s isa HipBuf ? (d .= s) : _upload!(d, s)
    end
    dsts
end
function OpenADMIXTURE.copyto_sync!(dsts::Vector{<:AbstractArray}, srcs::Vector{<:HipBuf})
    length(dsts) == length(srcs) || throw(ArgumentError("copyto_sync!: lengths differ"))
    for (d, s) in zip(dsts, srcs)
        d isa HipBuf ? (d .= s) : _download!(d, s)
    end
    dsts
end
function OpenADMIXTURE.copyto_sync!(dsts::Vector{<:HipBuf}, srcs::Vector{<:HipBuf})
    length(dsts) == length(srcs) || throw(ArgumentError("copyto_sync!: lengths differ"))
    for (d, s) in zip(dsts, srcs)
        d .= s
    end
    dsts
end

# ---- the context: CuAdmixData's counterpart (cuda/structs.jl:1-15) --------------------------------------

mutable struct HipAdmixData
    h::Ptr{Cvoid}
    I::Int
    J::Int
    K::Int
    q::HipBuf{2}
    q_next::HipBuf{2}
    p::HipBuf{2}
    p_next::HipBuf{2}
    XtX_q::HipBuf{3}
    Xtz_q::HipBuf{2}
    XtX_p::HipBuf{3}
    Xtz_p::HipBuf{2}
    doublemissing::Vector{Int64}
end

function HipAdmixData(h::Ptr{Cvoid}, I::Int, J::Int, K::Int)
    dm = zeros(Int64, I)
    hk_check(h, ccall((:hk_double_missing, libhk), Cint, (Ptr{Cvoid}, Ptr{Int64}), h, dm))
    dh = HipAdmixData(h, I, J, K,
                      HipBuf{2}(h, HK_BUF_Q, (K, I)), HipBuf{2}(h, HK_BUF_Q_NEXT, (K, I)),
                      HipBuf{2}(h, HK_BUF_P, (K, 4J)), HipBuf{2}(h, HK_BUF_P_NEXT, (K, 4J)),
                      HipBuf{3}(h, HK_BUF_XTX_Q, (K, K, I)), HipBuf{2}(h, HK_BUF_XTZ_Q, (K, I)),
                      HipBuf{3}(h, HK_BUF_XTX_P, (K, K, 4J)), HipBuf{2}(h, HK_BUF_XTZ_P, (K, 4J)), dm)
    finalizer(x -> (x.h != C_NULL && ccall((:hk_ctx_destroy, libhk), Cint, (Ptr{Cvoid},), x.h); x.h = C_NULL), dh)
    dh
end

"""
    _cu_admixture_base(d, g_la, I, J) -> (d_cu, nothing)

driver.jl:153-158 with use_gpu=true: upload the first I individuals and 2J SNPs of the SnpArray's payload into a
libhaplo_hip context (packed once on the device), count double-missing pairs over the whole file as
AdmixData2 does (structs.jl:169), and set q, q_next, p, p_next from d as CuAdmixData does (cuda/structs.jl:24-25).
T2 = Float32 selects the MIXED engine (FP32 per-pair products, FP64 sums); set HAPLO_HIP_T2=reference to get the
reference's FP32 accumulators instead.
"""
function _cu_admixture_base(d::AdmixData2{T,T2}, g_la::SnpLinAlg{T2}, I::Int, J::Int) where {T,T2}
    g = g_la.s.data                                   # UInt8, ceil(I_file/4) x nSNP, one column per SNP
    prec = T2 == Float32 ? (get(ENV, "HAPLO_HIP_T2", "") == "reference" ? HK_PREC_FP32 : HK_PREC_MIXED) :
           HK_PREC_FP64
    r = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:hk_ctx_create, libhk), Cint, (Ref{Ptr{Cvoid}}, Int64, Int64, Cint, Ptr{UInt8}, Int64, Int64, Cint),
               r, I, J, d.K, g, size(g, 1), size(g, 2), prec)
    rc == 0 || error("hk_ctx_create: ", unsafe_string(ccall((:hk_create_error, libhk), Cstring, ())))
    h = r[]
    hk_check(h, ccall((:hk_count_double_missing_file, libhk), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Int64, Int64),
                      h, g, size(g, 1), size(g, 2)))
    dh = HipAdmixData(h, I, J, d.K)
    OpenADMIXTURE.copyto_sync!([dh.q, dh.q_next, dh.p, dh.p_next], [d.q_T2, d.q_T2, d.p_T2, d.p_T2])
    dh, nothing
end

# ---- runners (cuda/runners.jl) ----------------------------------------------------------------------------

function loglikelihood(d::HipAdmixData, ::Nothing, q::HipBuf{2}, p::HipBuf{2})
    ll = Ref{Float64}(0.0)
    hk_check(d.h, ccall((:hk_loglik_buf, libhk), Cint, (Ptr{Cvoid}, Cint, Cint, Ref{Float64}), d.h, q.id, p.id, ll))
    ll[]
end

function em!(d::HipAdmixData, ::Nothing)
    hk_check(d.h, ccall((:hk_em_buf, libhk), Cint, (Ptr{Cvoid}, Cint), d.h, 0))
    d.q_next, d.p_next
end

function em_p!(d::HipAdmixData, ::Nothing)
    hk_check(d.h, ccall((:hk_em_buf, libhk), Cint, (Ptr{Cvoid}, Cint), d.h, 1))
    d.p_next
end

update_q_cuda!(d::HipAdmixData, ::Nothing) =
    (hk_check(d.h, ccall((:hk_update_q_buf, libhk), Cint, (Ptr{Cvoid},), d.h)); nothing)
update_p_cuda!(d::HipAdmixData, ::Nothing) =
    (hk_check(d.h, ccall((:hk_update_p_buf, libhk), Cint, (Ptr{Cvoid},), d.h)); nothing)

# ---- fit level: init_em! + admixture_qn! on the device -----------------------------------------------------

struct HkFitOpts                                      # hk_fit_opts
    em_iters::Cint; n_iter::Cint; rtol::Cdouble; Qn::Cint; fix_q::Cint; fix_p::Cint
    q_penalty::Cint; p_penalty::Cint; rho::Cdouble; eps::Cdouble; verbose::Cint; profile::Cint
    use_graph::Cint; em_fix_q::Cint; guarantee_increase::Cint; qn_mode::Cint; batch_rounds::Cint
end

struct HkFitResult                                    # hk_fit_result
    ll::Cdouble; ll_init::Cdouble; iters::Cint; converged::Cint; qp_failures::Cint; seconds::Cdouble
    t_qpass_ms::Cdouble; t_ppass_ms::Cdouble; t_ll_ms::Cdouble; t_qstep_ms::Cdouble; t_pstep_ms::Cdouble
    t_em_ms::Cdouble; t_qn_ms::Cdouble
    n_qpass::Cint; n_ppass::Cint; n_ll::Cint; n_qstep::Cint; n_pstep::Cint; n_em::Cint; n_qn::Cint
    qp_iter_limit::Cint
end

"""
    hip_fit!(d, dh; em_iters=5, n_iter=1000, rtol=1e-7, Q=3, ...)

init_em! + admixture_qn! (algorithms_outer.jl:1-172) on the device from d.q, d.p; leaves the fit in d.
"""
function hip_fit!(d::AdmixData2, dh::HipAdmixData; em_iters=5, n_iter=1000, rtol=1e-7, Q=3, fix_q=false,
                  fix_p=false, guarantee_increase=false, em_fix_q=false, mode=:ZAL, ll_start=NaN, verbose=false)
    hk_check(dh.h, ccall((:hk_set_params, libhk), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                         dh.h, Array{Float64}(d.q), Array{Float64}(d.p)))
    o = Ref(HkFitOpts(em_iters, n_iter, rtol, Q, fix_q, fix_p, 0, 1, 1e7, 1e-5, verbose, 0, 1, em_fix_q,
                      guarantee_increase, mode == :LBQN ? 1 : 0, 1))
    res = Ref{HkFitResult}()
    trace = zeros(Float64, max(n_iter, 1))
    hk_check(dh.h, ccall((:hk_fit, libhk), Cint, (Ptr{Cvoid}, Ref{HkFitOpts}, Cdouble, Ptr{Float64}, Ref{HkFitResult}),
                         dh.h, o, ll_start, trace, res))
    q = zeros(Float64, dh.K, dh.I)
    p = zeros(Float64, dh.K, 4dh.J)
    hk_check(dh.h, ccall((:hk_get_params, libhk), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), dh.h, q, p))
    d.q .= q
    d.p .= p
    d.ll_new = res[].ll
    d
end

# ---- additions beyond the reference: model selection, learned init_q, SKFR on the device -------------------

"""
    hip_select(dh, Ks, restarts; rtol=1e-7, seed=1, precision=HK_PREC_MIXED) -> Vector of (K, ll, aic, iters)

Fits of K in Ks x `restarts` random starts over dh's genotype upload (hk_ctx_create_shared), run concurrently by
hk_fit_batch with per-K autotuned pass shapes; AIC(K) = -2 ll + 2 (I (K - 1) + 3 K J).
"""
function hip_select(dh::HipAdmixData, Ks, restarts; rtol=1e-7, seed=1, precision=HK_PREC_MIXED, n_iter=1000)
    rng = Random.MersenneTwister(seed)
    ctxs, specs = Ptr{Cvoid}[], Tuple{Int,Int}[]
    try
        for K in Ks
            qc, pc, qs, ps = Ref{Cint}(0), Ref{Cint}(0), Ref{Cint}(0), Ref{Cint}(0)
            for r in 1:restarts
                h = Ref{Ptr{Cvoid}}(C_NULL)
                hk_check(dh.h, ccall((:hk_ctx_create_shared, libhk), Cint, (Ref{Ptr{Cvoid}}, Ptr{Cvoid}, Cint, Cint),
                                     h, dh.h, K, precision))
                push!(ctxs, h[]); push!(specs, (K, r))
                q = rand(rng, K, dh.I); q ./= sum(q, dims=1)          # AdmixData2's random start (structs.jl:155-199)
                p = rand(rng, K, 4dh.J)
                for j in 1:dh.J, k in 1:K
                    s = sum(@view p[k, 4(j-1)+1:4j]); p[k, 4(j-1)+1:4j] ./= s
                end
                hk_check(h[], ccall((:hk_set_params, libhk), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}), h[], q, p))
                if r == 1
                    hk_check(h[], ccall((:hk_autotune, libhk), Cint, (Ptr{Cvoid}, Cint, Ref{Cint}, Ref{Cint}), h[], 2, qc, pc))
                    hk_check(h[], ccall((:hk_get_pass_shape, libhk), Cint, (Ptr{Cvoid}, Ref{Cint}, Ref{Cint}), h[], qs, ps))
                else
                    hk_check(h[], ccall((:hk_set_tuning, libhk), Cint, (Ptr{Cvoid}, Cint, Cint), h[], qc[], pc[]))
                    hk_check(h[], ccall((:hk_set_pass_shape, libhk), Cint, (Ptr{Cvoid}, Cint, Cint), h[], qs[], ps[]))
                end
            end
        end
        n = length(ctxs)
        opts = [HkFitOpts(5, n_iter, rtol, 3, 0, 0, 0, 1, 1e7, 1e-5, 0, 0, 1, 0, 0, 0, 1) for _ in 1:n]
        res = Vector{HkFitResult}(undef, n)
        fin = zeros(Float64, n)
        rc = ccall((:hk_fit_batch, libhk), Cint, (Ptr{Ptr{Cvoid}}, Cint, Ptr{HkFitOpts}, Ptr{Float64},
                   Ptr{HkFitResult}, Ptr{Float64}), ctxs, n, opts, C_NULL, res, fin)
        rc == 0 || error("hk_fit_batch: ", join(hk_error.(ctxs), "; "))
        [(K = K, restart = r, ll = x.ll, aic = -2x.ll + 2(dh.I * (K - 1) + 3K * dh.J), iters = x.iters)
         for ((K, r), x) in zip(specs, res)]
    finally
        foreach(h -> ccall((:hk_ctx_destroy, libhk), Cint, (Ptr{Cvoid},), h), ctxs)
    end
end

"""
    hip_init_q(dh, mlp; n_blocks=min(dh.J, 20000), seed=77) -> K x I Matrix{Float64}

The learned init_q (driver.jl:56,184-185): the network's softmax over K from the device-side projection of the
first n_blocks SNP pairs (hk_mlp_predict), floored at 1e-5 and renormalised.  mlp comes from hk_mlp_create +
hk_mlp_set_params with the shipped weights.
"""
function hip_init_q(dh::HipAdmixData, mlp::Ptr{Cvoid}; n_blocks=min(dh.J, 20000), seed=77)
    P = zeros(Float32, dh.K, dh.I)                    # row-major I x K on the C side = K x I column-major here
    hk_check(dh.h, ccall((:hk_mlp_predict, libhk), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, UInt64, Ptr{Float32}),
                         mlp, dh.h, n_blocks, seed, P))
    Q = max.(Float64.(P), 1e-5)
    Q ./= sum(Q, dims=1)
end

"""
    hip_skfr(dh, K, sparsity; mode=0, max_iter=50, tries=1, seed=1) -> (clusters, aims)

_filter_SKFR's sparse k-means (driver.jl:84-106) on the device over dh's blocks: mode 0 = sparsekmeans1
(:global), 1 = sparsekmeans2 (:local); tries > 1 = sparsekmeans_repeat.  Returns 1-based labels and SNP indices,
most important block first.
"""
function hip_skfr(dh::HipAdmixData, K, sparsity; mode=0, max_iter=50, tries=1, seed=UInt64(1))
    clusters = zeros(Int32, dh.I)
    aims = zeros(Int32, 2 * min(dh.J, K * sparsity))
    n, wss, iters = Ref{Cint}(0), Ref{Float64}(0), Ref{Cint}(0)
    hk_check(dh.h, ccall((:hk_skfr, libhk), Cint,
                         (Ptr{Cvoid}, Cint, Cint, Cint, Cint, Cint, UInt64, Ptr{Int32}, Ptr{Int32}, Ptr{Int32},
                          Ref{Cint}, Ref{Float64}, Ref{Cint}, Ptr{Float64}),
                         dh.h, K, sparsity, mode, max_iter, tries, seed, C_NULL, clusters, aims, n, wss, iters, C_NULL))
    clusters .+ 1, aims[1:n[]] .+ 1
end
