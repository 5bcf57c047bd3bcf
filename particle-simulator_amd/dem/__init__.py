"""MI355X-native drop-in for the reference's ``dem`` package hot path.

Module paths, class names, constructor keywords and method signatures follow
the reference (dem/force_calculator, dem/octree, dem/collision_detector,
dem/solver, dem/solution, dem/particle_state, ...); the compute behind the
force, convection, contact and octree operators is hand-written HIP for
gfx950 reached through ``dem._native`` (libdemhip.so).
"""
