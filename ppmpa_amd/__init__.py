"""ppmpa_amd — MI355X (gfx950) implementation of the PPM-PA hot path of
eijian/ppmpa: photon emission and bouncing, photon-map build, radius query and
radiance gather, eye rays, and the PPM-PA pass schedule with FP64
accumulation, as hand-written HIP kernels behind the C ABI of include/ppm.h.
"""
from . import _abi
from .scene import (Camera, Light, Material, Mesh, Scene, Shape, CONFIGS, two_spheres_scene, caustic_scene, cornell_box,
                    batch_scenes, icosphere, inverse_albedos, inverse_scene, mesh_scene,
                    glass, lambertian, mirror, parallelogram, parallelogram_light, pinhole, plain, point_light,
                    polygon, read_camera, reference_room, simple, sphere, sun_light, ts)
from .tracer import (Batch, Context, DeviceScene, PhotonMap, PhotonSet, PpmError, Session, build_photonmap, pass_cfg,
                     photon_power, read_map, run_passes, trace_classic, trace_eye, trace_photons)

__all__ = [n for n in dir() if not n.startswith("_")]
