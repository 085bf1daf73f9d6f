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

    out.update(complete_sail_models())

    os.makedirs(os.path.join(HERE, "setups"), exist_ok=True)
    for name, text in out.items():
        json.loads(text)
        with open(os.path.join(HERE, "setups", name + ".json"), "w") as f:
            f.write(text)
        print("wrote", name)


def complete_sail_models():
    """CompleteSailModelBuilder strings (lifting_line/complete_sail_model.py) built the way the
    reference's sail-comparison example does
    (interfaces/pystormbird/examples/6 - simple sail comparison/compare_sail_types.py:75-100)."""
    from stormbird_setup.spatial_vector import SpatialVector
    from stormbird_setup.lifting_line import (SimulationBuilder, QuasiSteadySettings, QuasiSteadyWakeSettings,
                                              SymmetryCondition, CompleteSailModelBuilder)
    from stormbird_setup.lifting_line.simulation_builder import DynamicSettings
    from stormbird_setup.lifting_line.wake import DynamicWakeBuilder
    from stormbird_setup.simplified_setup.simple_sail_setup import SimpleSailSetup, SailType
    from stormbird_setup.controller import (ControllerBuilder, ControllerSetPoints, FlowMeasurementSettings,
                                            MeasurementSettings, MeasurementType)
    from stormbird_setup.wind import WindEnvironment, InflowCorrections, InflowCorrectionsSingleDirection
    from stormbird_setup.range import Range

    out = {}
    # one single-element wing sail, quasi-steady, z symmetry (compare_sail_types.py:66-100)
    sim = SimulationBuilder(simulation_settings=QuasiSteadySettings(
        wake=QuasiSteadyWakeSettings(symmetry_condition=SymmetryCondition.Z)))
    sail = SimpleSailSetup(position=SpatialVector(x=0.0, y=0.0, z=0.0), chord_length=5.0, height=35.0,
                           sail_type=SailType.WingSailSingleElement)
    sim.line_force_model.add_wing_builder(sail.wing_builder())
    out["csm_wing_sail_quasi_steady"] = CompleteSailModelBuilder(
        lifting_line_simulation=sim, controller=ControllerBuilder(set_points=[sail.controller_set_points()]),
        wind_environment=WindEnvironment()).to_json_string()

    # three wing sails with a shed wake, inflow corrections and an effective-angle controller
    dt, chord, velocity = 0.5, 4.0, 10.0
    sim = SimulationBuilder(simulation_settings=DynamicSettings(
        wake=DynamicWakeBuilder.new_default(time_step=dt, chord_length=chord, velocity=velocity)))
    sim.simulation_settings.wake.nr_panels_per_line_element = 24
    sim.simulation_settings.wake.symmetry_condition = SymmetryCondition.Z
    sim.line_force_model.nr_sections = 12
    set_points = []
    for k, (x, y) in enumerate(((-20.0, 0.0), (0.0, 3.0), (20.0, 0.0))):
        sail = SimpleSailSetup(position=SpatialVector(x=x, y=y, z=4.0), chord_length=chord, height=24.0,
                               sail_type=SailType.WingSailSingleElement)
        sim.line_force_model.add_wing_builder(sail.wing_builder())
        sp = ControllerSetPoints.new_default_wing_sail_single_element(max_angle_deg=12.0 + k)
        sp.use_effective_angle_of_attack = k == 1
        sp.max_local_wing_angle_change_rate = 0.2 if k == 2 else None
        set_points.append(sp)
    n = 12
    heights = [4.0 + 24.0 * (i + 0.5) / n for i in range(n)]
    corrections = []
    for d, (mag, ang) in enumerate(((0.9, 0.05), (1.1, -0.08))):
        corrections.append(InflowCorrectionsSingleDirection(
            height_values=heights * 3,
            magnitude_corrections=[mag + 0.01 * i for i in range(3 * n)],
            angle_corrections=[ang * (1.0 - i / (3.0 * n)) for i in range(3 * n)],
            wing_indices=[Range(start=w * n, end=(w + 1) * n) for w in range(3)]))
    env = WindEnvironment(water_plane_height=1.0, inflow_corrections=InflowCorrections(
        apparent_wind_directions=[-0.5, 0.7], corrections=corrections))
    controller = ControllerBuilder(
        set_points=set_points,
        flow_measurement_settings=FlowMeasurementSettings(
            angle_of_attack=MeasurementSettings(measurement_type=MeasurementType.Max, start_index=2, end_offset=3),
            wind_direction=MeasurementSettings(start_index=0, end_offset=0),
            wind_velocity=MeasurementSettings(measurement_type=MeasurementType.Min)),
        use_input_velocity_for_apparent_wind_direction=True)
    out["csm_three_sails_dynamic_inflow"] = CompleteSailModelBuilder(
        lifting_line_simulation=sim, controller=controller, wind_environment=env).to_json_string()
    return out


def wind_profiles(reference_root):
    """Velocity profiles evaluated by the reference's own Python models
    (stormbird_setup/wind/velocity_variation/power_model.py:44-54 and logarithmic_model.py:209-262,
    the same expressions as velocity_variation/power_model.rs:28-40 and logarithmic_model.rs:26-80),
    written to tests/golden/wind_profiles.json."""
    from stormbird_setup.wind import PowerModel, LogarithmicModel
    heights = [0.0, 0.5, 1.0, 2.0, 5.0, 10.0, 17.5, 30.0, 55.0, 100.0]
    cases = []
    for kw in ({"reference_velocity": 8.0}, {"reference_velocity": 12.0, "reference_height": 20.0, "power_factor": 0.14},
               {"reference_velocity": 5.0, "power_factor": 0.0}):
        m = PowerModel(**kw)
        cases.append({"PowerModel": m.model_dump(), "heights": heights,
                      "velocity": [float(m.velocity_at_height(h)) for h in heights]})
    for kw in ({"friction_velocity": 0.4, "surface_roughness": 0.0002},
               {"friction_velocity": 0.6, "surface_roughness": 0.001, "obukhov_length": 80.0},
               {"friction_velocity": 0.3, "surface_roughness": 0.0005, "obukhov_length": -40.0},
               {"friction_velocity": 0.5, "surface_roughness": 0.01, "obukhov_length": -500.0,
                "unstable_coefficient": 16.0, "von_karman_constant": 0.4}):
        m = LogarithmicModel(**kw)
        cases.append({"LogarithmicModel": m.model_dump(), "heights": heights,
                      "velocity": [float(m.velocity_at_height(h)) for h in heights],
                      "businger_dyer_unscaled": [float(m.businger_dyer_unscaled_correction(h)) for h in heights]})
    with open(os.path.join(HERE, "wind_profiles.json"), "w") as f:
        json.dump(cases, f, indent=1)
    print("wrote wind_profiles")


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    main(root)
    wind_profiles(root)
