"""
RadiK (MI355X build): the RadiK.jl API over libradik_hip.so.

Drop-in for the reference's exports (RadiK.jl src/RadiK.jl:20-23) with the
same names, signatures, Val switches and [k, num_tasks] output layout; the
work runs in hand-written gfx950 kernels behind a C ABI (include/radik.h).
Arrays are AMDGPU.jl ROCArrays; the calls run on AMDGPU's task-local stream
and synchronise before returning, as the reference does (algorithm.jl:353).

This file is not exercised in this repository's test suite (no Julia
toolchain); the Python mirror in radik.jl_amd/radik binds the same symbols.
"""
module RadiK

using AMDGPU

export topk_radix_select!, topk, topk!, RadiKWorkspace

const libradik = normpath(joinpath(@__DIR__, "..", "..", "..", "lib", "libradik_hip.so"))

const LARGEST_F = UInt32(0x01)
const ASCEND_F = UInt32(0x02)
const WITHSCALE_F = UInt32(0x04)
const WITHIDXIN_F = UInt32(0x08)
const WITHPACKING_F = UInt32(0x10)

last_error() = unsafe_string(ccall((:radik_last_error, libradik), Cstring, ()))

function check(rc::Cint)
    rc == 0 && return nothing
    error(last_error())            # ErrorException, as the reference raises
end

stream_ptr() = AMDGPU.stream().stream

# RadiKWorkspace(backend, n, num_tasks, ::Type{I}=Int32)       algorithm.jl:58-83
mutable struct RadiKWorkspace{I}
    handle::Ptr{Cvoid}
    n::Int
    num_tasks::Int
end

function RadiKWorkspace(backend, n, num_tasks, ::Type{I}=Int32) where {I}
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:radik_ws_create, libradik), Cint,
                (Ptr{Ptr{Cvoid}}, Int64, Int64, Cint, Ptr{Cvoid}),
                h, n, num_tasks, sizeof(I), stream_ptr()))
    ws = RadiKWorkspace{I}(h[], Int(n), Int(num_tasks))
    finalizer(ws) do w
        w.handle == C_NULL || ccall((:radik_ws_destroy, libradik), Cint, (Ptr{Cvoid},), w.handle)
        w.handle = C_NULL
    end
    return ws
end

# topk_radix_select!(...)                                       algorithm.jl:162-368
function topk_radix_select!(
    val_out::AbstractArray{Float32},
    idx_out::AbstractArray{IdxT},
    ws::RadiKWorkspace{I},
    val_in::AbstractArray{Float32},
    idx_in::AbstractArray{IdxT},
    task_lens::AbstractVector{<:Integer},
    k::Integer,
    ::Val{LARGEST}=Val(true),
    ::Val{ASCEND}=Val(true),
    ::Val{WITHSCALE}=Val(false),
    ::Val{WITHIDXIN}=Val(false),
    ::Val{WITHPACKING}=Val(true),
    ::Val{VLOAD_SIZE_BYTES}=Val(32)
) where {IdxT, I, LARGEST, ASCEND, WITHSCALE, WITHIDXIN, WITHPACKING, VLOAD_SIZE_BYTES}
    sizeof(IdxT) in (4, 8) || error("index type must be 32- or 64-bit, got $IdxT")
    flags = (LARGEST ? LARGEST_F : 0x0) | (ASCEND ? ASCEND_F : 0x0) | (WITHSCALE ? WITHSCALE_F : 0x0) |
            (WITHIDXIN ? WITHIDXIN_F : 0x0) | (WITHPACKING ? WITHPACKING_F : 0x0)
    lens = Int64.(collect(task_lens))
    iptr = WITHIDXIN ? pointer(idx_in) : Ptr{IdxT}(C_NULL)
    check(ccall((:radik_topk_radix_select, libradik), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Ptr{Cvoid}, Cint, Ptr{Float32}, Ptr{Cvoid}, Ptr{Int64}, Int64, Int64,
                 UInt32, Cint, Ptr{Cvoid}),
                ws.handle, pointer(val_out), pointer(idx_out), sizeof(IdxT), pointer(val_in), iptr, lens,
                length(lens), k, flags, VLOAD_SIZE_BYTES, stream_ptr()))
    return val_out, idx_out
end

# topk!(val_out, idx_out, data, task_lens, k; indices, largest, rev)  algorithm.jl:477-522
function topk!(
    val_out::AbstractArray{Float32, 2},
    idx_out::AbstractArray{IdxT, 2},
    data::AbstractArray{Float32},
    task_lens::AbstractVector{<:Integer},
    k::Integer;
    indices::Union{Nothing, AbstractArray{IdxT}}=nothing,
    largest=true,
    rev=false
) where IdxT
    num_tasks = length(task_lens)
    expected_size = (k, num_tasks)
    size(val_out) == expected_size || error("val_out must have size $expected_size, got $(size(val_out))")
    size(idx_out) == expected_size || error("idx_out must have size $expected_size, got $(size(idx_out))")
    if indices !== nothing && !isempty(indices)
        length(indices) == length(data) || error(
            "indices length must equal data length. " *
            "Got $(length(indices)) vs $(length(data))")
    end
    withidx = indices !== nothing && !isempty(indices)
    ws = RadiKWorkspace(nothing, maximum(task_lens), num_tasks)
    iarr = withidx ? indices : idx_out
    topk_radix_select!(val_out, idx_out, ws, data, iarr, task_lens, k,
                       Val(largest), Val(!rev), Val(false), Val(withidx), Val(true))
end

topk!(val_out, idx_out, data, k; indices=nothing, largest=true, rev=false) =
    topk!(val_out, idx_out, data, [length(data)], k, indices=indices, largest=largest, rev=rev)

# topk(data, task_lens, k; ...)                                 algorithm.jl:528-554
function topk(data::AbstractArray{Float32}, task_lens::AbstractVector{<:Integer}, k::Integer;
              indices::Union{Nothing, AbstractArray}=nothing, largest=true, rev=false)
    num_tasks = length(task_lens)
    val_out = AMDGPU.zeros(Float32, k, num_tasks)
    idx_type = indices === nothing ? Int32 : eltype(indices)
    idx_out = AMDGPU.zeros(idx_type, k, num_tasks)
    topk!(val_out, idx_out, data, task_lens, k, indices=indices, largest=largest, rev=rev)
    return val_out, idx_out
end

# topk(data, k; ...)                                            algorithm.jl:556-559
function topk(data, k; indices=nothing, largest=true, rev=false)
    val_out, idx_out = topk(data, [length(data)], k; indices=indices, largest=largest, rev=rev)
    return val_out[:, 1], idx_out[:, 1]
end

end # module RadiK
