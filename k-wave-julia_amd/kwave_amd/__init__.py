"""kwave_amd -- MI355X-native drop-in for KWave.jl's kspace_first_order.

Host-side mirror of the reference API (KWave.jl/src/{grid,medium,source,
sensor}.jl and solver/first_order.jl) over the C ABI of libkwave_hip.so
(include/kwave_hip.h).  The time loop runs only on the GPU; if the shared
library or the GPU is missing, the solver raises.
"""
from .model import (AbstractKWaveGrid, Additive, AdditiveNoCorrection, Dirichlet, ElasticMedium, ElasticSource,
                    KWaveGrid, KWaveGrid1D,
                    KWaveGrid2D, KWaveGrid3D, KWaveMedium, KWaveSensor, KWaveSource, SimulationOutput,
                    SourceMode, VALID_RECORD_FIELDS, has_p0, has_pressure_source, has_velocity_source,
                    is_homogeneous, is_lossless, is_nonlinear, is_time_reversal, make_time,
                    validate_record_fields)
from .builders import make_arc, make_ball, make_bowl, make_circle, make_disc, tone_burst
from .karray import (ArcElement, BowlElement, DiscElement, KWaveArray, RectElement, add_arc_element,
                     add_bowl_element, add_disc_element, add_rect_element, combine_sensor_data,
                     get_array_binary_mask, get_distributed_source_signal, get_element_binary_mask)
from .solver import (Simulation, build_desc, kspace_first_order, kspace_first_order_as, kspace_first_order_firings,
                     quantize_map)
from .batch import Batch, kspace_first_order_batch
from .adjoint import gradient_c0, misfit_rrule
from .elastic import ElasticSimulation, pstd_elastic_2d, pstd_elastic_3d
from .transducer import (KWaveTransducer, combine_transducer_sensor_data, get_transducer_binary_mask,
                         get_transducer_source)
from .hdf5 import read_matrix, read_output, write_attributes, write_flags, write_grid, write_matrix
from . import cli, hdf5, tuning
from ._lib import KWaveHIPError, LIB_PATH

__all__ = [
    "AbstractKWaveGrid", "ElasticMedium", "ElasticSource", "ElasticSimulation", "pstd_elastic_2d",
    "pstd_elastic_3d", "KWaveTransducer", "get_transducer_binary_mask", "get_transducer_source",
    "combine_transducer_sensor_data", "KWaveGrid", "KWaveGrid1D", "KWaveGrid2D", "KWaveGrid3D", "make_time",
    "KWaveMedium", "KWaveSource", "KWaveSensor", "SimulationOutput", "SourceMode", "Dirichlet",
    "Additive", "AdditiveNoCorrection", "VALID_RECORD_FIELDS", "validate_record_fields",
    "is_homogeneous", "is_lossless", "is_nonlinear", "is_time_reversal", "has_p0",
    "has_pressure_source", "has_velocity_source", "kspace_first_order", "kspace_first_order_as", "kspace_first_order_firings", "Simulation", "build_desc", "quantize_map",
    "Batch", "kspace_first_order_batch", "gradient_c0", "misfit_rrule", "tuning", "tone_burst", "make_disc", "make_circle", "make_arc", "make_ball", "make_bowl",
    "KWaveArray", "ArcElement", "BowlElement", "DiscElement", "RectElement", "add_arc_element", "add_bowl_element",
    "add_disc_element", "add_rect_element", "get_element_binary_mask", "get_array_binary_mask",
    "get_distributed_source_signal", "combine_sensor_data", "KWaveHIPError", "LIB_PATH",
    "write_matrix", "write_grid", "write_flags", "write_attributes", "read_matrix", "read_output", "hdf5", "cli",
]
