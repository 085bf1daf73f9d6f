"""Writes the setup-string fixtures in tests/golden/setups/ with the reference's own setup
package (interfaces/stormbird_setup, pydantic models whose JSON the Rust SimulationBuilder
reads).  The strings are what a user of the reference feeds to Simulation::new_from_string; the
tests load every one unchanged.  Run in a container that holds the reference:

    python tests/golden/make_setup_fixtures.py /root/reference
"""
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(reference_root):
    sys.path.insert(0, os.path.join(reference_root, "interfaces", "stormbird_setup", "src"))
    from stormbird_setup.spatial_vector import SpatialVector
    from stormbird_setup.section_models import SectionModel, Foil
    from stormbird_setup.line_force_model import LineForceModelBuilder, WingBuilder
    from stormbird_setup.lifting_line.simulation_builder import (SimulationBuilder, DynamicSettings,
                                                                 QuasiSteadySettings)
    from stormbird_setup.lifting_line.wake import (DynamicWakeBuilder, QuasiSteadyWakeSettings,
                                                   ViscousCoreLength, ViscousCoreLengthEvolution,
                                                   SymmetryCondition, FirstWakePointsDirection)
    from stormbird_setup.lifting_line.solver import SimpleIterative, Linearized
    from stormbird_setup.lifting_line.velocity_corrections import VelocityCorrections

    out = {}

    def heaving_wing(dynamic, nr_sections=32, span=32.0, chord=1.0, velocity=8.0):
        # interfaces/pystormbird/examples/4 - moving wing/heaving_wing.py:70-105
        dt = 0.1 * chord / velocity
        wing = WingBuilder(section_points=[SpatialVector(z=-span / 2.0), SpatialVector(z=span / 2.0)],
                           chord_vectors=[SpatialVector(x=chord), SpatialVector(x=chord)],
                           section_model=SectionModel(model=Foil()))
        lfm = LineForceModelBuilder(nr_sections=nr_sections)
        lfm.add_wing_builder(wing)
        if dynamic:
            settings = DynamicSettings(wake=DynamicWakeBuilder.new_default(time_step=dt, chord_length=chord,
                                                                           velocity=velocity))
        else:
            settings = QuasiSteadySettings()
        return SimulationBuilder(line_force_model=lfm, simulation_settings=settings).to_json_string()

    out["heaving_wing_dynamic"] = heaving_wing(True)
    out["heaving_wing_quasi_steady"] = heaving_wing(False)

    # several sails, every setting away from its default
    lfm = LineForceModelBuilder(nr_sections=12, density=1.2,
                                rotation=SpatialVector(z=0.1), translation=SpatialVector(x=2.0, y=-1.0))
    for k, x in enumerate((-10.0, 10.0)):
        lfm.add_wing_builder(WingBuilder(
            section_points=[SpatialVector(x=x), SpatialVector(x=x, z=8.0), SpatialVector(x=x, z=20.0)],
            chord_vectors=[SpatialVector(x=6.0), SpatialVector(x=5.0), SpatialVector(x=3.0)],
            section_model=SectionModel(model=Foil(cl_zero_angle=0.2 * k, cd_min=0.012,
                                                  mean_positive_stall_angle=math.radians(18.0))),
            nr_sections=10 + 4 * k))
    lfm.local_wing_angles = [0.05, -0.1]
    wake = DynamicWakeBuilder(nr_panels_per_line_element=40,
                              viscous_core_length=ViscousCoreLength.new_absolute(0.3),
                              viscous_core_length_evolution=ViscousCoreLengthEvolution.new_sin_increase(
                                  last_panel_value_absolute=1.5, evolution_length_factor=0.5),
                              symmetry_condition=SymmetryCondition.Z,
                              first_panel_relative_length=0.3, last_panel_relative_length=0.5,
                              first_wake_points_direction=FirstWakePointsDirection.Freestream,
                              ratio_of_wake_affected_by_induced_velocities=0.5,
                              shape_damping_factor=0.1)
    solver = SimpleIterative(max_iterations_per_time_step=60, damping_factor=0.08,
                             velocity_corrections=VelocityCorrections.new_max_induced_velocity_magnitude(0.5),
                             start_with_linearized_solution=True)
    out["two_sails_dynamic_all_settings"] = SimulationBuilder(
        line_force_model=lfm, simulation_settings=DynamicSettings(solver=solver, wake=wake)).to_json_string()

    qs = QuasiSteadySettings(solver=Linearized(velocity_corrections=VelocityCorrections.new_fixed_magnitude_equal_to_freestream(),
                                               disable_viscous_corrections=False),
                             wake=QuasiSteadyWakeSettings(wake_length_factor=50.0,
                                                          symmetry_condition=SymmetryCondition.Z,
                                                          viscous_core_length=ViscousCoreLength.new_absolute(0.2)))
    out["two_sails_quasi_steady_linearized"] = SimulationBuilder(line_force_model=lfm,
                                                                 simulation_settings=qs).to_json_string()
    out["two_sails_quasi_steady_iterative"] = SimulationBuilder(
        line_force_model=lfm, simulation_settings=QuasiSteadySettings.new_steady_iterative()).to_json_string()

    os.makedirs(os.path.join(HERE, "setups"), exist_ok=True)
    for name, text in out.items():
        json.loads(text)
        with open(os.path.join(HERE, "setups", name + ".json"), "w") as f:
            f.write(text)
        print("wrote", name)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
