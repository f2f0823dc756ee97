"""CPU oracle for the differentiable substep.  TEST INFRASTRUCTURE ONLY.

Nothing in the product package imports this directory.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
use it, and only as the checker (or as the timed CPU baseline), never as the
thing measured on the GPU.

The oracle restates, on the CPU, the arithmetic of the reference kernels:

* ``mpm.py``     -- MLS-MPM object path (compute_new_F, svd, p2g elastic /
                    von Mises / multi-material, check_grid_occupy, grid_op,
                    g2p) following ``difftactile/object_model/mpm_elastic.py``,
                    ``mpm_plastic.py`` and ``multi_obj.py``;
* ``fem.py``     -- FEM pad (init_pos, update, update2 dome + gripper
                    variants) following ``difftactile/sensor_model/fem_sensor.py``
                    and ``gripper_fem.py``;
* ``contact.py`` -- grid penalty contact (find_closest, find_sdf,
                    calculate_contact_force, check_collision, collision,
                    clamp_grid) following ``difftactile/tasks/grasp_plastic.py``;
* ``rope.py``, ``coupled.py``, ``gripper_scene.py`` -- the PBD cable and its
                    pad contact, the object-pad substep and the gripper
                    scene's composition;
* ``kinematics.py``, ``dome_pose.py`` -- the jaw chain and the dome pad's pose
                    control; ``markers.py`` -- markers and the task losses;
                    ``deform.py`` -- the grid-mass deformation losses;
* ``cpu_mpm.cpp`` (``cpu_mpm.py``, ``build_cpu.py``) -- a C++17/OpenMP
                    float32 restatement of one MPM substep and its adjoint in
                    the reference's loop structure, checked against
                    ``mpm.py`` and timed as the CPU baseline.

Forward code is written in torch on the CPU so that reverse-mode autograd over
it reproduces what Taichi's reverse-mode AD (``kernel.grad``) does for the
same code: the exact derivative of the written arithmetic, with no gradient
through integer casts or branch conditions.  The one hand-written adjoint of
the reference, ``svd_grad`` (``mpm_elastic.py:144-173``), replaces autograd's
SVD backward, exactly as the reference calls ``svd_grad(f)`` in place of
``svd.grad`` (``grasp_plastic.py:177``).

How each function is pinned (beyond restating the reference) is listed in the
module docstrings and exercised by ``tests/test_oracle_*.py``: closed forms,
conservation identities, FP64 central differences, ``np.linalg.svd`` and
brute force.  Functions with no such check say "parity unpinned".
"""
