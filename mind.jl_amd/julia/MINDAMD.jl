"""
    MINDAMD

Stdlib-only Julia binding of `libmind`, the MI355X build of MIND.jl's simulation
step. MIND.jl keeps its model-building API (`MicroModel`, `add_population!`,
`insert!`, `add_pointprocess!`, `register_spike_source!`, `connect!`, the macro
`MacroModel`/`roi`/`use_macro!`/...); its `build!`, `initialize!`, `run!` and
`step!` (src/model/build.jl:120-220, src/runtime/initialize.jl:256-429,
src/runtime/step.jl:4-86) flatten the model into `ModelArrays` and call the
functions below instead of allocating CUDA.jl state.

Every C struct mirrors `include/mind.h` field for field; every call checks the
returned code and raises `MindError` with `mind_last_error()`'s text, as the
reference raises `ErrorException` for device error words (src/runtime/utils/step.jl:190-196).

Julia is not part of the build image this binding was written in, so it has not
been executed; the Python layer (`mind_amd`) makes the same calls and is tested.
"""
module MINDAMD

export ModelArrays, MindError, build_state, initialize_state!, run_state!, step_state!,
       destroy_state!, read_voltage, voltage_record_spec, field_record_spec, read_value,
       connection_weights, inject_event!, inject_spike!

const libmind = get(ENV, "MIND_LIBRARY", joinpath(@__DIR__, "..", "csrc", "libmind.so"))

struct MindError <: Exception
    code::Int32
    message::String
end
Base.showerror(io::IO, e::MindError) = print(io, "libmind error ", e.code, ": ", e.message)

function check(code::Integer)
    code == 0 && return nothing
    msg = unsafe_string(ccall((:mind_last_error, libmind), Cstring, ()))
    throw(MindError(Int32(code), msg))
end

# ---- C structs (include/mind.h) ------------------------------------------------
struct ModelDesc
    celsius::Float64
    n_morphologies::Int32
    morph_offsets::Ptr{Int64}
    sec_parent::Ptr{Int32}
    sec_parent_x::Ptr{Float64}
    sec_child_x::Ptr{Float64}
    sec_nseg::Ptr{Int32}
    sec_length::Ptr{Float64}
    sec_diameter::Ptr{Float64}
    sec_Ra::Ptr{Float64}
    sec_cm::Ptr{Float64}
    sec_point_offsets::Ptr{Int64}
    points::Ptr{Float64}
    n_cells::Int32
    cell_morphology::Ptr{Int32}
    cell_gid::Ptr{Int64}
    cell_voltage::Ptr{Float64}
    n_density::Int64
    dens_cell::Ptr{Int32}
    dens_section::Ptr{Int32}
    dens_mech::Ptr{Int32}
    dens_param_offsets::Ptr{Int64}
    n_points::Int64
    pt_mech::Ptr{Int32}
    pt_cell::Ptr{Int32}
    pt_section::Ptr{Int32}
    pt_x::Ptr{Float64}
    pt_param_offsets::Ptr{Int64}
    param_field::Ptr{Int32}
    param_value::Ptr{Float64}
    n_sources::Int64
    src_kind::Ptr{Int32}
    src_cell::Ptr{Int32}
    src_section::Ptr{Int32}
    src_x::Ptr{Float64}
    src_threshold::Ptr{Float64}
    src_point::Ptr{Int32}
    src_sid::Ptr{Int64}
    n_connections::Int64
    con_source::Ptr{Int32}
    con_target::Ptr{Int32}
    con_weight::Ptr{Float64}
    con_delay::Ptr{Float64}
    pt_spike_offsets::Ptr{Int64}   # PatternStim spike times (CSR over points), or NULL
    pt_spike_times::Ptr{Float64}
end

struct BuildOptions
    dt::Float64
    seed::UInt64
    method::Int32          # 0 plain, 1 batched
    cells_per_group::Int32 # 0 = autotune
    window_steps::Int32    # 0 = longest the minimum delay allows
    event_capacity::Int64  # 0 = automatic
    device::Int32
    secondorder::Int32     # 0 backward Euler, 1 or 2 Crank-Nicolson
end

struct RecordSpec
    kind::Int32   # 0 voltage at (cell, section, x); 1 point-process field; 2 density field
    cell::Int32
    section::Int32
    x::Float64
    point::Int32
    mech::Int32
    field::Int32
end

struct ResultC
    n_samples::Int64
    time::Ptr{Float64}
    n_records::Int64
    traces::Ptr{Float64}
    n_spikes::Int64
    spike_times::Ptr{Float64}
    spike_sids::Ptr{Int64}
    run_seconds::Float64
end
ResultC() = ResultC(0, C_NULL, 0, C_NULL, 0, C_NULL, C_NULL, 0.0)

"""
    ModelArrays

The flattened `MicroModel` (0-based indices, as in `include/mind.h`). MIND.jl's
`build!` fills it from its morphologies, populations, insertions, point
processes, sources and connections in creation order; the vectors must stay
alive until `build_state` returns (the library copies them).
"""
Base.@kwdef mutable struct ModelArrays
    celsius::Float64 = 6.3
    morph_offsets::Vector{Int64} = Int64[0]
    sec_parent::Vector{Int32} = Int32[]
    sec_parent_x::Vector{Float64} = Float64[]
    sec_child_x::Vector{Float64} = Float64[]
    sec_nseg::Vector{Int32} = Int32[]
    sec_length::Vector{Float64} = Float64[]
    sec_diameter::Vector{Float64} = Float64[]
    sec_Ra::Vector{Float64} = Float64[]
    sec_cm::Vector{Float64} = Float64[]
    cell_morphology::Vector{Int32} = Int32[]
    cell_gid::Vector{Int64} = Int64[]
    cell_voltage::Vector{Float64} = Float64[]
    dens_cell::Vector{Int32} = Int32[]
    dens_section::Vector{Int32} = Int32[]
    dens_mech::Vector{Int32} = Int32[]
    dens_param_offsets::Vector{Int64} = Int64[0]
    pt_mech::Vector{Int32} = Int32[]
    pt_cell::Vector{Int32} = Int32[]
    pt_section::Vector{Int32} = Int32[]
    pt_x::Vector{Float64} = Float64[]
    pt_param_offsets::Vector{Int64} = Int64[0]
    param_field::Vector{Int32} = Int32[]
    param_value::Vector{Float64} = Float64[]
    src_kind::Vector{Int32} = Int32[]
    src_cell::Vector{Int32} = Int32[]
    src_section::Vector{Int32} = Int32[]
    src_x::Vector{Float64} = Float64[]
    src_threshold::Vector{Float64} = Float64[]
    src_point::Vector{Int32} = Int32[]
    src_sid::Vector{Int64} = Int64[]
    con_source::Vector{Int32} = Int32[]
    con_target::Vector{Int32} = Int32[]
    con_weight::Vector{Float64} = Float64[]
    con_delay::Vector{Float64} = Float64[]
    pt_spike_offsets::Vector{Int64} = Int64[]   # empty, or one offset per point + 1
    pt_spike_times::Vector{Float64} = Float64[]
end

function descriptor(a::ModelArrays)
    p(v) = isempty(v) ? C_NULL : pointer(v)
    ModelDesc(a.celsius, Int32(length(a.morph_offsets) - 1), p(a.morph_offsets), p(a.sec_parent),
              p(a.sec_parent_x), p(a.sec_child_x), p(a.sec_nseg), p(a.sec_length), p(a.sec_diameter),
              p(a.sec_Ra), p(a.sec_cm), C_NULL, C_NULL, Int32(length(a.cell_morphology)),
              p(a.cell_morphology), p(a.cell_gid), p(a.cell_voltage), length(a.dens_cell), p(a.dens_cell),
              p(a.dens_section), p(a.dens_mech), p(a.dens_param_offsets), length(a.pt_mech), p(a.pt_mech),
              p(a.pt_cell), p(a.pt_section), p(a.pt_x), p(a.pt_param_offsets), p(a.param_field),
              p(a.param_value), length(a.src_kind), p(a.src_kind), p(a.src_cell), p(a.src_section),
              p(a.src_x), p(a.src_threshold), p(a.src_point), p(a.src_sid), length(a.con_source),
              p(a.con_source), p(a.con_target), p(a.con_weight), p(a.con_delay),
              p(a.pt_spike_offsets), p(a.pt_spike_times))
end

# ---- calls -----------------------------------------------------------------------
"""
    build_state(arrays; dt_micro=0.025, seed=1, method=:batched, cells_per_group=0,
                window_steps=0, event_capacity=0, device=0, secondorder=0) -> Ptr{Cvoid}

`build!` (src/model/build.jl:120-131): packs and uploads the model; the batched
method autotunes its cells-per-workgroup plan.
"""
function build_state(a::ModelArrays; dt_micro = 0.025, seed = 1, method = :batched,
                     cells_per_group = 0, window_steps = 0, event_capacity = 0, device = 0,
                     secondorder = 0)
    secondorder in (0, 1, 2) || throw(ArgumentError("secondorder must be 0, 1 or 2"))
    opts = Ref(BuildOptions(dt_micro, UInt64(seed), method === :batched ? Int32(1) : Int32(0),
                            Int32(cells_per_group), Int32(window_steps), Int64(event_capacity),
                            Int32(device), Int32(secondorder)))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve a begin
        desc = Ref(descriptor(a))
        check(ccall((:mind_build, libmind), Cint, (Ptr{ModelDesc}, Ptr{BuildOptions}, Ptr{Ptr{Cvoid}}),
                    desc, opts, out))
    end
    return out[]
end

"`initialize!` (src/runtime/initialize.jl:256-429)."
initialize_state!(s::Ptr{Cvoid}) = check(ccall((:mind_initialize, libmind), Cint, (Ptr{Cvoid},), s))

"`step!` (src/runtime/utils/step.jl:215): one micro step."
step_state!(s::Ptr{Cvoid}) = check(ccall((:mind_step, libmind), Cint, (Ptr{Cvoid},), s))

destroy_state!(s::Ptr{Cvoid}) = ccall((:mind_state_destroy, libmind), Cvoid, (Ptr{Cvoid},), s)

voltage_record_spec(cell, section, x) = RecordSpec(0, cell, section, x, -1, -1, -1)
field_record_spec(point, field) = RecordSpec(1, -1, -1, 0.0, point, -1, field)

"""
    run_state!(s, stop; records=RecordSpec[], record_spikes=true)

`run!` (src/runtime/step.jl:4-74): advances to `stop` (an integer number of steps
ahead) and returns `(time, traces, spike_times, spike_sids, run_seconds)` with
traces as a `n_records × n_samples` matrix, like `SimulationResult`
(src/runtime/recording.jl:78-87).
"""
function run_state!(s::Ptr{Cvoid}, stop::Real; records::Vector{RecordSpec} = RecordSpec[],
                    record_spikes::Bool = true)
    res = Ref(ResultC())
    check(ccall((:mind_run, libmind), Cint,
                (Ptr{Cvoid}, Float64, Ptr{RecordSpec}, Int64, Int32, Ptr{ResultC}),
                s, Float64(stop), isempty(records) ? C_NULL : records, length(records),
                Int32(record_spikes), res))
    r = res[]
    try
        time = copy(unsafe_wrap(Array, r.time, r.n_samples))
        traces = r.n_records == 0 ? zeros(0, r.n_samples) :
                 permutedims(copy(unsafe_wrap(Array, r.traces, (r.n_samples, r.n_records))))
        st = r.n_spikes == 0 ? Float64[] : copy(unsafe_wrap(Array, r.spike_times, r.n_spikes))
        sids = r.n_spikes == 0 ? Int64[] : copy(unsafe_wrap(Array, r.spike_sids, r.n_spikes))
        return (; time, traces, spike_times = st, spike_sids = sids, run_seconds = r.run_seconds)
    finally
        ccall((:mind_result_free, libmind), Cvoid, (Ptr{ResultC},), res)
    end
end

"`read_value` (src/runtime/state.jl:586-600): the current value of one record selector."
function read_value(s::Ptr{Cvoid}, spec::RecordSpec)
    out = Ref{Float64}(0.0)
    check(ccall((:mind_read_value, libmind), Cint, (Ptr{Cvoid}, Ref{RecordSpec}, Ptr{Float64}), s, spec, out))
    return out[]
end

"`connection_weights` (src/runtime/state.jl:603-616) of the 1-based connection number."
function connection_weights(s::Ptr{Cvoid}, connection::Integer)
    out = zeros(8)
    n = Ref{Int32}(0)
    check(ccall((:mind_connection_weights, libmind), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Int32, Ptr{Int32}),
                s, Int64(connection) - 1, out, Int32(length(out)), n))
    return out[1:n[]]
end

"`inject_event!` (src/runtime/state.jl:618-636): 1-based connection, absolute arrival time."
inject_event!(s::Ptr{Cvoid}, connection::Integer, time::Real) =
    check(ccall((:mind_inject_event, libmind), Cint, (Ptr{Cvoid}, Int64, Float64), s, Int64(connection) - 1,
                Float64(time)))

"`inject_spike!` (src/runtime/state.jl:638-659): 1-based external source index, absolute time."
inject_spike!(s::Ptr{Cvoid}, source::Integer, time::Real) =
    check(ccall((:mind_inject_spike, libmind), Cint, (Ptr{Cvoid}, Int64, Float64), s, Int64(source) - 1,
                Float64(time)))

"Membrane voltage of every node in canonical node order."
function read_voltage(s::Ptr{Cvoid})
    nodes, comps = Ref{Int64}(0), Ref{Int64}(0)
    check(ccall((:mind_node_count, libmind), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}), s, nodes, comps))
    v = Vector{Float64}(undef, nodes[])
    check(ccall((:mind_read_voltage, libmind), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64), s, v, length(v)))
    return v
end

end # module
