"""Host layer of the MI355X-native MIND simulation step.

Mirrors the reference's model-building API (MicroModel, Section,
add_population, insert, add_pointprocess, register_spike_source, connect,
build, initialize, run, step, *_record) over libmind's C ABI.
"""
from .mechanisms import Mechanism, builtin, compile_mod, load_mod  # noqa: F401
from .model import (  # noqa: F401
    Cell, CellSection, IonSpecies, IonStyle, MicroModel, Morphology, PointProcess, Section, SimulationResult,
    SpikeSource, default_ions, ion_style,
    add_pointprocess, add_population, batch_plan, build, checkpoint, connect, connection_weights, destroy,
    restore,
    global_record, initialize, inject_event, inject_spike, insert, ion_record, read_value, section, section_group,
    set_batch_plan, set_global,
    mechanism_record, read_field, read_voltage, register_spike_source, run, step,
    voltage_record, window_steps,
)
from .swc import load_swc  # noqa: F401,E402
from .asc import load_asc  # noqa: F401,E402
