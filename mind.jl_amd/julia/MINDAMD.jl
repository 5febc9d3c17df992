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
       connection_weights, inject_event!, inject_spike!, Mechanism, compile_mod, load_mod,
       builtin, column, globals, flatten, MacroArrays, flatten_macro, build_macro_state, run_macro_state!,
       SurrogateArrays, build_surrogate, initialize_surrogate!, run_surrogate!, destroy_surrogate!

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
    n_species::Int32
    species_name::Ptr{Cstring}
    species_valence::Ptr{Float64}
    species_internal::Ptr{Float64}
    species_external::Ptr{Float64}
    species_reversal::Ptr{Float64}
    species_nernst::Ptr{Int32}
    n_ion_settings::Int64
    ionset_cell::Ptr{Int32}
    ionset_section::Ptr{Int32}
    ionset_species::Ptr{Int32}
    ionset_what::Ptr{Int32}
    ionset_value::Ptr{Float64}
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
    sec_point_offsets::Vector{Int64} = Int64[]  # empty, or one offset per section + 1 into points
    points::Vector{Float64} = Float64[]         # x, y, z, diameter per pt3d sample
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
    # ion species (empty: default_ions) and section ion settings (include/mind.h)
    species_name::Vector{String} = String[]
    species_valence::Vector{Float64} = Float64[]
    species_internal::Vector{Float64} = Float64[]
    species_external::Vector{Float64} = Float64[]
    species_reversal::Vector{Float64} = Float64[]
    species_nernst::Vector{Int32} = Int32[]
    ionset_cell::Vector{Int32} = Int32[]
    ionset_section::Vector{Int32} = Int32[]
    ionset_species::Vector{Int32} = Int32[]
    ionset_what::Vector{Int32} = Int32[]
    ionset_value::Vector{Float64} = Float64[]
    species_ptrs::Vector{Ptr{UInt8}} = Ptr{UInt8}[]  # filled by descriptor
end

function descriptor(a::ModelArrays)
    p(v) = isempty(v) ? C_NULL : pointer(v)
    ModelDesc(a.celsius, Int32(length(a.morph_offsets) - 1), p(a.morph_offsets), p(a.sec_parent),
              p(a.sec_parent_x), p(a.sec_child_x), p(a.sec_nseg), p(a.sec_length), p(a.sec_diameter),
              p(a.sec_Ra), p(a.sec_cm), p(a.sec_point_offsets), p(a.points), Int32(length(a.cell_morphology)),
              p(a.cell_morphology), p(a.cell_gid), p(a.cell_voltage), length(a.dens_cell), p(a.dens_cell),
              p(a.dens_section), p(a.dens_mech), p(a.dens_param_offsets), length(a.pt_mech), p(a.pt_mech),
              p(a.pt_cell), p(a.pt_section), p(a.pt_x), p(a.pt_param_offsets), p(a.param_field),
              p(a.param_value), length(a.src_kind), p(a.src_kind), p(a.src_cell), p(a.src_section),
              p(a.src_x), p(a.src_threshold), p(a.src_point), p(a.src_sid), length(a.con_source),
              p(a.con_source), p(a.con_target), p(a.con_weight), p(a.con_delay),
              p(a.pt_spike_offsets), p(a.pt_spike_times),
              Int32(length(a.species_name)), Ptr{Cstring}(p(species_pointers!(a))), p(a.species_valence),
              p(a.species_internal), p(a.species_external), p(a.species_reversal), p(a.species_nernst),
              length(a.ionset_cell), p(a.ionset_cell), p(a.ionset_section), p(a.ionset_species),
              p(a.ionset_what), p(a.ionset_value))
end

function species_pointers!(a::ModelArrays)
    a.species_ptrs = [pointer(n) for n in a.species_name]
    return a.species_ptrs
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

# ---- mechanisms (compile_mod / load_mod / builtin, src/compiler/frontend.jl:58-110) ----
struct MechDesc
    name::Cstring
    filename::Cstring
    kind::Int32
    n_fields::Int32
    has_net_receive::Int32
    has_state::Int32
    compiled::Int32
    digest::UInt64
end

"""
    Mechanism

A mechanism slot of libmind: a builtin (ids 0-9, `include/mind.h`) or a MOD text
lowered at run time by `compile_mod`. `fields` are the column names in column
order (element k of array field `a` is `"a[k]"`), `defaults` their values.
"""
struct Mechanism
    id::Int32
    name::String
    kind::Symbol        # :density, :point or :artificial
    fields::Vector{String}
    defaults::Vector{Float64}
    has_net_receive::Bool
    compiled::Bool
end

function Mechanism(id::Integer)
    d = Ref{MechDesc}()
    check(ccall((:mind_mech_info, libmind), Cint, (Int32, Ptr{MechDesc}), id, d))
    n = Int(d[].n_fields)
    names = Vector{Cstring}(undef, max(n, 1))
    vals = zeros(Float64, max(n, 1))
    count = Ref{Int32}(0)
    check(ccall((:mind_mech_fields, libmind), Cint, (Int32, Ptr{Cstring}, Ptr{Float64}, Int32, Ptr{Int32}),
                id, names, vals, Int32(n), count))
    kinds = (:density, :point, :artificial)
    return Mechanism(Int32(id), unsafe_string(d[].name), kinds[d[].kind + 1],
                     [unsafe_string(names[k]) for k in 1:n], vals[1:n], d[].has_net_receive != 0,
                     d[].compiled != 0)
end

"""
    compile_mod(source; filename="<string>") -> Mechanism

`compile_mod` (frontend.jl:74-110): lowers MOD text in libmind (C++ NMODL
frontend) into a mechanism slot; the device phases are compiled with hipRTC when
a model that uses it is built. The same text and filename give the same slot.
Malformed or unsupported input raises `MindError` (code 1) naming file and line.
"""
function compile_mod(source::AbstractString; filename::AbstractString = "<string>")
    text = String(source)
    id = Ref{Int32}(-1)
    check(ccall((:mind_compile_mod, libmind), Cint, (Ptr{UInt8}, Csize_t, Cstring, Ptr{Int32}),
                text, sizeof(text), String(filename), id))
    return Mechanism(id[])
end

"`load_mod(path)` (frontend.jl:74-80): `compile_mod` of a MOD file."
load_mod(path::AbstractString) = compile_mod(read(path, String); filename = path)

"`builtin(name)` (frontend.jl:58-72): one of the ten packaged mechanisms."
function builtin(name)
    for id in 0:9
        m = Mechanism(id)
        m.name == String(name) && return m
    end
    throw(ArgumentError("unknown mechanism $name; load its MOD with load_mod"))
end

"""
    column(m::Mechanism, field, index=1) -> Int32

0-based column of `field` (element `index`, 1-based, of an array field).
"""
function column(m::Mechanism, field, index::Integer = 1)
    name = String(field)
    k = findfirst(==(name), m.fields)
    if k !== nothing
        index == 1 || throw(ArgumentError("$(m.name).$name is not an array field"))
        return Int32(k - 1)
    end
    k = findfirst(==("$name[$index]"), m.fields)
    k === nothing && throw(ArgumentError("mechanism $(m.name) has no field $name (element $index)"))
    return Int32(k - 1)
end

"""
    globals(m::Mechanism) -> Vector{String}

The GLOBAL columns of a mechanism (`mind_mech_globals`): one value every instance
shares (global_ref, src/model/references.jl:147-154). `compile_mod` refuses a MOD
that assigns one; `flatten`'s `globals` keyword sets their values for every
instance. The builtins declare none.
"""
function globals(m::Mechanism)
    n = length(m.fields)
    flags = zeros(UInt8, max(n, 1))
    count = Ref{Int32}(0)
    check(ccall((:mind_mech_globals, libmind), Cint, (Int32, Ptr{UInt8}, Int32, Ptr{Int32}),
                m.id, flags, Int32(n), count))
    return [m.fields[k] for k in 1:n if flags[k] != 0]
end

# ---- flattening a MIND.jl model ----------------------------------------------------
"""
    flatten(model; mechanisms=Dict{Symbol, Mechanism}(), macro_inputs=false) -> ModelArrays

Flattens a MIND.jl `MicroModel` (src/model/micro/model.jl:15-50: morphologies of
`Section`s, `CellInstance`s with their `DensityBinding`s, `PointInstance`s,
`SpikeSourceDefinition`s and `Connection`s, all in creation order) into the
0-based arrays of `mind_model_desc`. A compiled mechanism resolves through
`mechanisms` (name => `Mechanism` from `compile_mod`/`load_mod` here), else to
the mechanism of the same name (`mind_mech_lookup`: the most recently compiled
one, else the builtin). Parameter values are numbers, or vectors for array
fields. `globals` (mechanism name => field => value) sets GLOBAL values for every
instance of a mechanism (the reference's global parameters, `pack_globals`,
src/model/build.jl:388). Connections from macro inputs (negative source index) are left out:
`flatten_macro` passes them as MACRO2MICRO bindings.

Raises `ArgumentError` for what the device path does not take: per-section
initial voltages, per-cell ion settings, ion species other than the defaults,
threshold sources on mechanism fields, NET_RECEIVE blocks with several
arguments, parameter values that are functions.
"""
function flatten(model; mechanisms::AbstractDict = Dict{Symbol, Mechanism}(), macro_inputs::Bool = false,
                 globals::AbstractDict = Dict{Symbol, Any}())
    a = ModelArrays(celsius = Float64(model.celsius))
    known = Dict{Symbol, Mechanism}(Symbol(k) => v for (k, v) in mechanisms)
    function resolve(compiled)
        name = Symbol(compiled.ir.name)
        haskey(known, name) && return known[name]
        id = Ref{Int32}(-1)
        check(ccall((:mind_mech_lookup, libmind), Cint, (Cstring, Ptr{Int32}), String(name), id))
        known[name] = Mechanism(id[])
        return known[name]
    end
    function params!(m::Mechanism, pairs)
        shared = get(globals, Symbol(m.name), Dict{Symbol, Any}())
        for (key, value) in Iterators.flatten((pairs, shared))
            if value isa Real
                push!(a.param_field, column(m, key))
                push!(a.param_value, Float64(value))
            elseif value isa AbstractVector{<:Real}
                for (k, x) in enumerate(value)
                    push!(a.param_field, column(m, key, k))
                    push!(a.param_value, Float64(x))
                end
            else
                throw(ArgumentError("parameter $key of $(m.name): only numbers and vectors of numbers"))
            end
        end
    end
    for (name, species) in model.ions  # default_ions (utils/biophysics.jl:31-42): fixed potentials
        defaults = Dict(:na => 50.0, :k => -77.0, :ca => 132.4579341637009)
        haskey(defaults, Symbol(name)) && species.reversal == defaults[Symbol(name)] && !species.nernst ||
            throw(ArgumentError("ion species $name differs from the default ions"))
    end
    # 1. morphologies: sections in order, parents by name
    section_index = Vector{Dict{Symbol, Int32}}()
    any_points = any(s -> length(s.points) >= 2, Iterators.flatten(model.morphologies))
    any_points && push!(a.sec_point_offsets, 0)
    for morphology in model.morphologies
        index = Dict{Symbol, Int32}(s.name => Int32(k - 1) for (k, s) in enumerate(morphology))
        push!(section_index, index)
        for s in morphology
            push!(a.sec_parent, s.parent === nothing ? Int32(-1) : index[s.parent])
            push!(a.sec_parent_x, s.parent_x)
            push!(a.sec_child_x, s.child_x)
            push!(a.sec_nseg, Int32(s.nseg))
            push!(a.sec_length, s.length)
            push!(a.sec_diameter, s.diameter)
            push!(a.sec_Ra, s.axial_resistivity)
            push!(a.sec_cm, s.capacitance)
            if any_points
                if length(s.points) >= 2
                    for q in s.points
                        append!(a.points, (q.x, q.y, q.z, q.diameter))
                    end
                end
                push!(a.sec_point_offsets, length(a.points) ÷ 4)
            end
        end
        push!(a.morph_offsets, length(a.sec_parent))
    end
    # 2. cells, then their density insertions in binding order
    for cell in model.cells
        isempty(cell.section_voltages) ||
            throw(ArgumentError("per-section initial voltages are outside the device path"))
        isempty(cell.ion_values) && isempty(cell.ion_styles) ||
            throw(ArgumentError("per-cell ion settings are outside the device path"))
        push!(a.cell_morphology, Int32(cell.morphology - 1))
        push!(a.cell_gid, Int64(cell.gid))
        push!(a.cell_voltage, Float64(cell.voltage))
    end
    for (c, cell) in enumerate(model.cells)
        index = section_index[cell.morphology]
        for binding in cell.bindings
            m = resolve(binding.mechanism)
            m.kind == :density || throw(ArgumentError("$(m.name) is not a density mechanism"))
            for name in binding.sections
                push!(a.dens_cell, Int32(c - 1))
                push!(a.dens_section, index[name])
                push!(a.dens_mech, m.id)
                params!(m, binding.parameters)
                push!(a.dens_param_offsets, length(a.param_field))
            end
        end
    end
    # 3. point processes; PatternStim spike times as a CSR over points
    spikes = any(p -> !(p.spike_times isa Tuple) && !isempty(p.spike_times), model.points)
    spikes && push!(a.pt_spike_offsets, 0)
    for p in model.points
        m = resolve(p.mechanism)
        push!(a.pt_mech, m.id)
        if p.cell == 0
            push!(a.pt_cell, Int32(-1))
            push!(a.pt_section, Int32(-1))
            push!(a.pt_x, 0.0)
        else
            push!(a.pt_cell, Int32(p.cell - 1))
            push!(a.pt_section, section_index[model.cells[p.cell].morphology][p.section])
            push!(a.pt_x, Float64(p.x))
        end
        params!(m, p.parameters)
        push!(a.pt_param_offsets, length(a.param_field))
        if spikes
            p.spike_times isa Tuple || append!(a.pt_spike_times, p.spike_times)
            push!(a.pt_spike_offsets, length(a.pt_spike_times))
        end
    end
    # 4. spike sources: voltage detectors, point-process outputs, external inputs
    for src in model.sources
        src.mechanism === nothing ||
            throw(ArgumentError("threshold sources on mechanism fields are outside the device path"))
        if src.point > 0
            kind, cell, sec, x = 1, -1, -1, 0.0
        elseif src.cell > 0
            kind, cell, sec, x = 0, src.cell - 1, section_index[model.cells[src.cell].morphology][src.section], src.x
        else
            kind, cell, sec, x = 2, -1, -1, 0.0
        end
        push!(a.src_kind, Int32(kind))
        push!(a.src_cell, Int32(cell))
        push!(a.src_section, Int32(sec))
        push!(a.src_x, Float64(x))
        push!(a.src_threshold, Float64(src.threshold))
        push!(a.src_point, Int32(src.point - 1))
        push!(a.src_sid, Int64(src.sid))
    end
    # 5. NetCons in creation order
    for con in model.connections
        if con.source_index <= 0
            macro_inputs || throw(ArgumentError("a macro input connection needs flatten_macro"))
            continue
        end
        con.weights === nothing && con.arity <= 1 ||
            throw(ArgumentError("NET_RECEIVE blocks with several arguments are outside the device path"))
        push!(a.con_source, Int32(con.source_index - 1))
        push!(a.con_target, Int32(con.target - 1))
        push!(a.con_weight, Float64(con.weight))
        push!(a.con_delay, Float64(con.delay))
    end
    return a
end

# ---- micro-macro coupling (mind_build_macro / mind_run_macro) ----------------------
struct MacroDesc
    n_roi::Int32
    dt_macro::Float64
    weights::Ptr{Float64}
    delays::Ptr{Float64}
    wc_params::Ptr{Float64}
    bin_ms::Ptr{Float64}
    input_offsets::Ptr{Int64}
    input_point::Ptr{Int32}
    input_weight::Ptr{Float64}
    input_delay::Ptr{Float64}
    output_offsets::Ptr{Int64}
    output_source::Ptr{Int32}
end

struct MacroResultC
    n_samples::Int64
    time::Ptr{Float64}
    n_roi::Int32
    traces::Ptr{Float64}
    n_spikes::Int64
    spike_times::Ptr{Float64}
    spike_sids::Ptr{Int64}
    run_seconds::Float64
end
MacroResultC() = MacroResultC(0, C_NULL, 0, C_NULL, 0, C_NULL, C_NULL, 0.0)

# Wilson-Cowan REGION parameters in WC_* order (include/mind.h) and their defaults
const WC_PARAMS = (:tau_e, :tau_i, :c_ee, :c_ei, :c_ie, :c_ii, :a_e, :theta_e, :a_i, :theta_i, :P, :Q,
                   :k_micro, :rate_scale, :E0, :I0)
const WC_DEFAULTS = (10.0, 10.0, 16.0, 12.0, 15.0, 3.0, 1.3, 4.0, 2.0, 3.7, 1.25, 0.0, 0.0, 100.0, 0.1, 0.05)

"The flattened ROI graph of a `MacroModel` (0-based, `mind_macro_desc` order)."
Base.@kwdef mutable struct MacroArrays
    dt_macro::Float64 = 0.1
    weights::Vector{Float64} = Float64[]     # [target * n + source]
    delays::Vector{Float64} = Float64[]
    wc_params::Vector{Float64} = Float64[]   # [roi * 16 + k]
    bin_ms::Vector{Float64} = Float64[]
    input_offsets::Vector{Int64} = Int64[0]
    input_point::Vector{Int32} = Int32[]
    input_weight::Vector{Float64} = Float64[]
    input_delay::Vector{Float64} = Float64[]
    output_offsets::Vector{Int64} = Int64[0]
    output_source::Vector{Int32} = Int32[]
end

function macro_descriptor(m::MacroArrays)
    p(v) = isempty(v) ? C_NULL : pointer(v)
    MacroDesc(Int32(length(m.bin_ms)), m.dt_macro, p(m.weights), p(m.delays), p(m.wc_params), p(m.bin_ms),
              p(m.input_offsets), p(m.input_point), p(m.input_weight), p(m.input_delay), p(m.output_offsets),
              p(m.output_source))
end

"""
    flatten_macro(model; dt_macro=0.1, mechanisms=Dict()) -> (ModelArrays, MacroArrays)

Flattens a MIND.jl `MacroModel` (src/model/macro/model.jl:55-162) whose ROIs
all run the Wilson-Cowan REGION (`use_macro!` parameters and initial state by
the names of `WC_PARAMS`; `E`/`I` as initial state), whose projections are
LinearDelay MACRO2MACRO rules (connectome weight and delay), whose MACRO2MICRO
bindings are PoissonDrive (weight and delay of their connection) and whose
MICRO2MACRO transforms are PopRate (`bin`, ms; every ROI with outputs shares one
bin). Other MIND-role mechanisms raise `ArgumentError`: the device macro step
runs these four.
"""
function flatten_macro(model; dt_macro::Real = 0.1, mechanisms::AbstractDict = Dict{Symbol, Mechanism}())
    model.micro === nothing && throw(ArgumentError("the macro model has no micro network"))
    micro = flatten(model.micro; mechanisms, macro_inputs = true)
    n = length(model.connectivity.labels)
    m = MacroArrays(dt_macro = Float64(dt_macro))
    W = zeros(n, n)
    for pr in model.projections
        Symbol(pr.mechanism.ir.name) == :LinearDelay ||
            throw(ArgumentError("MACRO2MACRO $(pr.mechanism.ir.name) is outside the device macro step"))
        W[pr.target, pr.source] = model.connectivity.weights[pr.target, pr.source]
    end
    for t in 1:n, s in 1:n
        push!(m.weights, W[t, s])
        push!(m.delays, model.connectivity.delays[t, s])
    end
    for r in 1:n
        region = model.regions[r]
        region === nothing && throw(ArgumentError("ROI $(model.connectivity.labels[r]) has no REGION owner"))
        Symbol(region.mechanism.ir.name) == :WilsonCowan ||
            throw(ArgumentError("REGION $(region.mechanism.ir.name) is outside the device macro step"))
        values = collect(WC_DEFAULTS)
        for (key, value) in region.parameters
            k = findfirst(==(Symbol(key)), WC_PARAMS)
            k === nothing && throw(ArgumentError("WilsonCowan has no parameter $key"))
            values[k] = Float64(value)
        end
        for (key, value) in region.initial_state
            values[findfirst(==(Symbol(key) == :E ? :E0 : :I0), WC_PARAMS)] = Float64(value)
        end
        append!(m.wc_params, values)
    end
    # MACRO2MICRO: the k-th input binding is the connection with source index -k
    edges = Dict(-con.source_index => con for con in model.micro.connections if con.source_index <= 0)
    order = sortperm([b.roi for b in model.inputs])  # stable: ROI, then creation order
    counts = zeros(Int64, n)
    for k in order
        b = model.inputs[k]
        Symbol(b.mechanism.ir.name) == :PoissonDrive ||
            throw(ArgumentError("MACRO2MICRO $(b.mechanism.ir.name) is outside the device macro step"))
        con = edges[k]
        push!(m.input_point, Int32(con.target - 1))
        push!(m.input_weight, Float64(con.weight))
        push!(m.input_delay, Float64(con.delay))
        counts[b.roi] += 1
    end
    m.input_offsets = vcat(Int64(0), cumsum(counts))
    # MICRO2MACRO: sources by SID, in source order
    bins = fill(1.0, n)
    sources = [Int32[] for _ in 1:n]
    for out in model.outputs
        Symbol(out.mechanism.ir.name) == :PopRate ||
            throw(ArgumentError("MICRO2MACRO $(out.mechanism.ir.name) is outside the device macro step"))
        bins[out.roi] = Float64(get(out.parameters, :bin, 1.0))
        for sid in out.sources
            push!(sources[out.roi], Int32(model.micro.source_indices[sid] - 1))
        end
    end
    for r in 1:n
        append!(m.output_source, sort(unique(sources[r])))
        push!(m.output_offsets, length(m.output_source))
    end
    m.bin_ms = bins
    return micro, m
end

"`build!` of a macro model (build.jl:120-220 with pack_macro and pack_transforms)."
function build_macro_state(micro::ModelArrays, m::MacroArrays; dt_micro = 0.025, seed = 1, method = :batched,
                           cells_per_group = 0, device = 0)
    opts = Ref(BuildOptions(dt_micro, UInt64(seed), method === :batched ? Int32(1) : Int32(0),
                            Int32(cells_per_group), Int32(0), Int64(0), Int32(device), Int32(0)))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve micro m begin
        d = Ref(descriptor(micro))
        md = Ref(macro_descriptor(m))
        check(ccall((:mind_build_macro, libmind), Cint,
                    (Ptr{ModelDesc}, Ptr{MacroDesc}, Ptr{BuildOptions}, Ptr{Ptr{Cvoid}}), d, md, opts, out))
    end
    return out[]
end

function macro_result(r::MacroResultC)
    ns, nr = Int(r.n_samples), Int(r.n_roi)
    time = copy(unsafe_wrap(Array, r.time, ns))
    # [5][n_roi][n_samples] row-major: E, I, rate (Hz), coupling, mean soma V
    raw = ns == 0 ? zeros(ns, nr, 5) : copy(unsafe_wrap(Array, r.traces, (ns, nr, 5)))
    names = (:E, :I, :rate, :coupling, :v)
    traces = NamedTuple{names}(Tuple(permutedims(raw[:, :, k]) for k in 1:5))
    st = r.n_spikes == 0 ? Float64[] : copy(unsafe_wrap(Array, r.spike_times, r.n_spikes))
    sids = r.n_spikes == 0 ? Int64[] : copy(unsafe_wrap(Array, r.spike_sids, r.n_spikes))
    return (; time, traces, spike_times = st, spike_sids = sids, run_seconds = r.run_seconds)
end

"""
    run_macro_state!(s, stop; record_spikes=true)

`run!` of a macro model (step.jl:4-74, advance_macro! step.jl:156-204): whole
macro steps to `stop`; traces are `n_roi × n_samples` matrices per quantity.
"""
function run_macro_state!(s::Ptr{Cvoid}, stop::Real; record_spikes::Bool = true)
    res = Ref(MacroResultC())
    check(ccall((:mind_run_macro, libmind), Cint, (Ptr{Cvoid}, Float64, Int32, Ptr{MacroResultC}),
                s, Float64(stop), Int32(record_spikes), res))
    try
        return macro_result(res[])
    finally
        ccall((:mind_macro_result_free, libmind), Cvoid, (Ptr{MacroResultC},), res)
    end
end

# ---- surrogate micro-macro coupling (no counterpart in the reference) ---------------
struct SurrogateDesc
    n_in::Int32
    hidden::Int32
    layers::Int32
    n_out::Int32
    history::Int32
    params::Ptr{Float32}
    n_params::Int64
    x_mean::Ptr{Float32}
    x_std::Ptr{Float32}
    y_mean::Ptr{Float32}
    y_std::Ptr{Float32}
    x_lo::Ptr{Float32}
    x_hi::Ptr{Float32}
    skip_w::Ptr{Float32}
    skip_b::Ptr{Float32}
    v_rest::Float64
end

"A trained surrogate network (`mind_surrogate_desc`): flat parameters and normalisation."
Base.@kwdef mutable struct SurrogateArrays
    hidden::Int32 = 256
    layers::Int32 = 3
    history::Int32 = 4
    params::Vector{Float32} = Float32[]
    x_mean::Vector{Float32} = Float32[]
    x_std::Vector{Float32} = Float32[]
    y_mean::Vector{Float32} = Float32[]
    y_std::Vector{Float32} = Float32[]
    x_lo::Vector{Float32} = Float32[]
    x_hi::Vector{Float32} = Float32[]
    skip_w::Vector{Float32} = Float32[]
    skip_b::Vector{Float32} = Float32[]
    v_rest::Float64 = -65.0
end

"""
    build_surrogate(m::MacroArrays, net::SurrogateArrays; device=0) -> Ptr{Cvoid}

The ROI graph of `m` with a trained MLP in place of every micro population
(`mind_build_surrogate`).
"""
function build_surrogate(m::MacroArrays, net::SurrogateArrays; device = 0)
    p(v) = isempty(v) ? Ptr{Float32}(C_NULL) : pointer(v)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve m net begin
        sd = Ref(SurrogateDesc(net.history + 2, net.hidden, net.layers, 2, net.history, p(net.params),
                               length(net.params), p(net.x_mean), p(net.x_std), p(net.y_mean), p(net.y_std),
                               p(net.x_lo), p(net.x_hi), p(net.skip_w), p(net.skip_b), net.v_rest))
        md = Ref(macro_descriptor(m))
        check(ccall((:mind_build_surrogate, libmind), Cint, (Ptr{MacroDesc}, Ptr{SurrogateDesc}, Int32,
                                                             Ptr{Ptr{Cvoid}}), md, sd, Int32(device), out))
    end
    return out[]
end

initialize_surrogate!(s::Ptr{Cvoid}) = check(ccall((:mind_initialize_surrogate, libmind), Cint, (Ptr{Cvoid},), s))
destroy_surrogate!(s::Ptr{Cvoid}) = ccall((:mind_surrogate_destroy, libmind), Cvoid, (Ptr{Cvoid},), s)

"""
    run_surrogate!(s, stop) -> (result, ood_windows, first_ood_time)

Whole macro steps to `stop`; `ood_windows` counts bins whose inputs left the
training range (the caller then runs the micro-macro model instead).
"""
function run_surrogate!(s::Ptr{Cvoid}, stop::Real)
    res = Ref(MacroResultC())
    ood = Ref{Int64}(0)
    first = Ref{Float64}(-1.0)
    check(ccall((:mind_run_surrogate, libmind), Cint, (Ptr{Cvoid}, Float64, Ptr{MacroResultC}, Ptr{Int64},
                                                        Ptr{Float64}), s, Float64(stop), res, ood, first))
    try
        return macro_result(res[]), ood[], first[]
    finally
        ccall((:mind_macro_result_free, libmind), Cvoid, (Ptr{MacroResultC},), res)
    end
end

end # module
