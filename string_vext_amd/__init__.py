"""MI355X (gfx950) guest-framework Lennard-Jones Vext.

A drop-in for the Vext path of String_Vext: the same Python API
(``cell``, ``input``, ``orientation``, ``physics``, ``neighbor``, ``native``,
``vext``) over libsv_vext.so, whose C ABI (include/sv_vext.h) also exports the
reference's ``evaluate_reused_neighbors``.  ``field`` holds the batched GPU
handles (plans, pose sets, batches); ``synthetic`` generates the seeded
benchmark frameworks.
"""

__version__ = "0.1.0"
