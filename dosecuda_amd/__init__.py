"""dosecuda_amd -- MI355X-native (gfx950, hand-written HIP) dose engine with
the call surface of DoseCUDA's ``dose_kernels`` extension.

* ``dosecuda_amd.dose_kernels``  drop-in module (proton_raytrace_cuda,
  proton_spot_cuda, photon_dose_cuda) over libdosehip.so's C ABI;
* ``dosecuda_amd.engine``        device-pointer API (torch tensors, streams);
* ``dosecuda_amd.beam_models``   loaders for the unchanged lookup tables;
* ``dosecuda_amd.plan``          dose grids, plans and beams with the reference's
  class names, computing plan sums on the device; ``dosecuda_amd.nrrd``,
  ``dosecuda_amd.nifti`` and ``dosecuda_amd.dicom`` read and write the
  NRRD, NIfTI and DICOM (CT, RT Dose, RT Structure Set, RT (Ion) Plan) files;
* ``dosecuda_amd.phantoms``      seeded synthetic phantoms and fields.
"""
__version__ = "0.1.0"
