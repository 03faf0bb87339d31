"""Drop-in projector operator: the omegatomo ``projectorClass`` surface on the HIP library.

Mirrors ``omegatomo.projector.proj.projectorClass`` (source/Python/omegatomo/projector/proj.py:36)
for the matrix-free projector path:

* configure attributes, then ``addProjector()`` (proj.py:520) and ``initProj()`` (proj.py:1913);
* ``forwardProject(f, subset)`` / ``backwardProject(y, subset)`` (proj.py:1923-1933);
* ``options * f`` and ``options.T() * y`` with ``options.subset = k`` (proj.py:1936-1951);
* the bookkeeping attributes ``N``, ``Nx/Ny/Nz`` (uint32 arrays), ``nMeas`` (cumulative subset
  starts), ``nMeasSubset``, ``nProjSubset``, ``nTotMeas``, ``index``, ``x0``.

Tensors are flat float32 CUDA (HIP) tensors in OMEGA's column-major order (x fastest; measurements
radial-bin fastest, then angle, then plane/projection), as with ``useTorch=True, useCUDA=True``
(projfunctions.py:182-185). Outputs are newly allocated per call. Work is queued on torch's
current stream; unlike the reference there is no device-wide synchronisation after each launch.

Coordinates: either the caller provides the projector's coordinate tables directly (``x`` and
``z``, as in source/Python/custom_detector_exampleSimple.py), or they are generated from
scanner parameters by :mod:`omega_amd.geometry` (a ring scanner sampled as parallel-chord
sinograms; a circular cone-beam trajectory for CT).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from . import geometry as G


def _stream_ptr(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Geometry:
    """Owns an omg_geom handle (device copies of the coordinate tables)."""

    def __init__(self, desc: _lib.GeomDesc, keep=()):
        self._keep = keep  # host arrays referenced by desc until create returns
        h = C.c_void_p()
        _lib.call("omg_geom_create", C.byref(desc), C.byref(h))
        self.handle = h
        self.desc = desc

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            _lib.load().omg_geom_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class projectorClass:
    """Projector operator with the omegatomo ``projectorClass`` interface (projector types 1, 2, 5)."""

    def __init__(self):
        # image
        self.Nx = 128
        self.Ny = 128
        self.Nz = 63
        self.FOVa_x = 0.0
        self.FOVa_y = 0.0
        self.axial_fov = 0.0
        self.oOffsetX = 0.0
        self.oOffsetY = 0.0
        self.oOffsetZ = 0.0
        # projector
        self.projector_type = 1
        self.tube_width_xy = 0.0
        self.tube_width_z = 0.0
        self.tube_radius = 1.0      # projector 3: radius of the tube of response (mm)
        self.voxel_radius = 1.0     # projector 3: voxel sphere radius relative to half a voxel's diagonal
        # data
        self.CT = False
        self.listmode = 0
        self.subsets = 1
        self.subsetType = 8
        self.seed = -1                  # subset type 3: numpy default_rng seed (< 0: unseeded)
        self.x = np.empty(0, dtype=np.float32)
        self.z = np.empty(0, dtype=np.float32)
        self.Ndist = 0
        self.Nang = 0
        self.NSinos = 0
        self.nRowsD = 0
        self.nColsD = 0
        self.nProjections = 0
        self.SinM = np.empty(0, dtype=np.float32)
        self.normalization_correction = False
        self.normalization = np.empty(0, dtype=np.float32)
        self.global_correction_factor = 1.0
        # multiplicative per-measurement correction (proj.py:88,295 -> -DSCATTER)
        self.additionalCorrection = False
        self.corrVector = np.empty(0, dtype=np.float32)
        # rays per LOR (proj.py:263-264); their spread is dPitchX (transaxial) / dPitchY (axial), which the
        # reference sets to cr_pz / cr_p for PET data (proj.py:937-939)
        self.n_rays_transaxial = 1
        self.n_rays_axial = 1
        self.cr_p = 0.0
        # time of flight (proj.py:81,166-168,332-333,455-456, 681-705)
        self.TOF_bins = 1
        self.TOF_bins_used = 1
        self.TOF_FWHM = 0.0             # timing resolution (s)
        self.TOF_width = 0.0            # bin width (s)
        self.TOF_offset = 0.0           # (s)
        self.TOFCenter = np.empty(0, dtype=np.float32)
        self.sigma_x = 0.0
        # attenuation correction (proj.py:85,212-214,225,296,300,358)
        self.attenuation_correction = False
        self.CTAttenuation = True       # True: vaimennus is an attenuation image (1/mm); False: one factor per measurement
        self.CT_attenuation = True
        self.vaimennus = np.empty(0, dtype=np.float32)
        self.attIncm = False            # the image is in 1/cm
        self.rotateAttImage = 0
        self.flipAttImageXY = False
        self.flipAttImageZ = False
        self.epps = 1e-5
        self.flat = 0.0                 # CT: blank-scan (flat-field) value of the measured counts (proj.py:359)
        # PET scanner (used when no coordinate tables are given)
        self.diameter = 0.0
        self.det_per_ring = 0
        self.rings = 0
        self.cr_pz = 0.0
        self.ring_difference = -1
        # OMEGA's block-ring scanner model (proj.py:127-169): used when these are set and x / z are not given
        self.blocks_per_ring = 0
        self.cryst_per_block = 0
        self.transaxial_multip = 1
        self.DOI = 0.0
        self.flip_image = False
        self.offangle = 0.0
        self.ndist_side = 1
        self.span = 3
        self.det_w_pseudo = 0            # > det_per_ring: a pseudo detector after every block (proj.py:220)
        self.ringGaps = np.empty(0, dtype=np.float32)   # axial gaps between blocks of rings (mm, proj.py:201)
        self.cryst_per_block_axial = 0   # rings per axial block (proj.py:129; default: all rings)
        self.nLayers = 1                 # crystal layers (only one is provided)
        # CT geometry (used when no coordinate tables are given)
        self.sourceToCRot = 0.0
        self.sourceToDetector = 0.0
        # interpolation projector (projector_type 4; proj.py:287): interpolation length in voxels of FOVa_x / Nx
        # (0: one voxel; negative: the Joseph variant, one sample per voxel plane)
        self.dL = 0.0
        self.dPitchX = 0.0
        self.dPitchY = 0.0
        self.angles = np.empty(0, dtype=np.float32)
        self.uV = np.empty(0, dtype=np.float32)
        self.pitch = False
        # backend switches kept for API compatibility
        self.useTorch = True
        self.useCUDA = True
        self.useCuPy = False
        self.x0 = np.empty(0, dtype=np.float32)
        # state
        self.subset = 0
        self.trans = False
        self.projectorAdded = False
        self.projectorInitialized = False
        self.nMultiVolumes = 0
        self._geom = None

    # ------------------------------------------------------------------ configuration
    def OMEGAErrorCheck(self):
        """Configuration checks with the reference's ValueError behaviour (proj.py:1068-1285)."""
        if self.FOVa_x > 0 and self.FOVa_y == 0:
            self.FOVa_y = self.FOVa_x
        if self.projector_type not in (1, 2, 3, 4, 5, 11, 22, 33, 44, 55):
            raise ValueError("Only projector types 1 (improved Siddon), 2 (orthogonal distance-based), 3 (volume "
                             "of intersection), 4 (interpolation-based) and 5 (branchless distance-driven) are "
                             "available in this implementation!")
        if self.projector_type in (5, 55) and not self.CT:
            raise ValueError("Projector type 5 is only supported with CT data!")
        if self.subsets > 1 and not self.listmode:
            if self.subsetType == 10:
                raise ValueError("Not supported in Python version!")
            if self.CT and self.subsetType not in (8, 9, 11):
                raise ValueError("Projection-level subset types 8, 9 and 11 are supported for CT data!")
            if self.subsetType not in (1, 2, 3, 4, 5, 8, 9, 11):
                raise ValueError("Subset types 1-5, 8, 9 and 11 are supported for sinogram data!")
        if self.subsets < 1:
            raise ValueError("Number of subsets must be at least 1!")
        if self.FOVa_x <= 0 or self.axial_fov <= 0:
            raise ValueError("FOVa_x and axial_fov must be positive!")
        if int(np.asarray(self.Nx).ravel()[0]) <= 0 or int(np.asarray(self.Ny).ravel()[0]) <= 0 \
                or int(np.asarray(self.Nz).ravel()[0]) <= 0:
            raise ValueError("Image dimensions must be positive!")
        if self.projector_type in (2, 22) and self.tube_width_xy <= 0 and self.tube_width_z <= 0:
            raise ValueError("Orthogonal distance-based projector needs tube_width_xy or tube_width_z > 0!")

    def _build_tables(self):
        if self.CT:
            if self.x.size == 0:
                if self.angles.size == 0 or self.sourceToCRot <= 0 or self.sourceToDetector <= 0:
                    raise ValueError("CT data needs either the x/z coordinate tables or angles, sourceToCRot and "
                                     "sourceToDetector!")
                ang = np.asarray(self.angles, dtype=np.float64).ravel()
                if np.max(np.abs(ang)) > 6.0 * math.pi:
                    ang = ang * (math.pi / 180.0)
                c, s = np.cos(ang), np.sin(ang)
                R, D = float(self.sourceToCRot), float(self.sourceToDetector)
                src = np.stack([R * c, R * s, np.zeros_like(c)], -1)
                det = np.stack([-(D - R) * c, -(D - R) * s, np.zeros_like(c)], -1)
                self.x = np.concatenate([src, det], -1).reshape(-1).astype(np.float32)
                self.z = np.stack([-s * self.dPitchX, c * self.dPitchX], -1).reshape(-1).astype(np.float32)
                self.nProjections = ang.size
            if self.nProjections == 0:
                self.nProjections = self.x.size // 6
            self.NSinos = self.nProjections
            self.Ndist, self.Nang = self.nRowsD, self.nColsD
        elif self.listmode:
            self.nRowsD, self.nColsD = 1, 1
            self.nProjections = self.x.size // 6
            self.NSinos = self.nProjections
        else:
            if self.x.size == 0 and self.blocks_per_ring > 0 and self.cryst_per_block > 0:
                self._omega_scanner_tables()
            if self.x.size == 0:
                if self.diameter <= 0 or self.det_per_ring <= 0:
                    raise ValueError("PET data needs either the x/z coordinate tables or diameter and det_per_ring!")
                R = self.diameter / 2.0
                if self.Nang == 0:
                    self.Nang = self.det_per_ring // 2
                if self.Ndist == 0:
                    self.Ndist = self.det_per_ring // 2
                self.x = G.ring_sinogram_xy(self.Ndist, self.Nang, math.pi * R / self.det_per_ring, R)
                if self.z.size == 0:
                    mrd = self.rings - 1 if self.ring_difference < 0 else self.ring_difference
                    self.z = G.ring_pair_planes_z(self.rings, self.cr_pz, mrd)
            if self.Ndist == 0 or self.Nang == 0:
                raise ValueError("Ndist and Nang are required with user-supplied coordinates!")
            self.nRowsD, self.nColsD = self.Ndist, self.Nang
            self.NSinos = self.z.size // 2
            self.nProjections = self.NSinos
        self.x = np.ascontiguousarray(self.x, dtype=np.float32).ravel()
        self.z = np.ascontiguousarray(self.z, dtype=np.float32).ravel()

    def _omega_scanner_tables(self):
        """Coordinate tables of OMEGA's block-ring scanner model from the scanner parameters (detcoord.py
        getCoordinates -> detectorCoordinates, sinogramCoordinates2D/3D; proj.py:654-679 for the derived
        sizes): blocks_per_ring x cryst_per_block crystals of pitch cr_p on a ring of diameter (+ 2 DOI),
        with a pseudo detector after every block when det_w_pseudo says so; rings of pitch cr_pz, with
        ringGaps between axial blocks of cryst_per_block_axial rings; Ndist x Nang sinograms, angles mashed
        when Nang < det_w_pseudo / 2; span-compressed ring pairs."""
        blocks = int(self.blocks_per_ring) * int(self.transaxial_multip)
        det = blocks * int(self.cryst_per_block)
        if self.det_per_ring == 0:
            self.det_per_ring = det
        if self.det_per_ring != det:
            raise ValueError("det_per_ring must equal blocks_per_ring * transaxial_multip * cryst_per_block")
        if self.det_w_pseudo == 0:
            self.det_w_pseudo = det
        pseudo = int(self.det_w_pseudo) > det
        if pseudo and int(self.det_w_pseudo) != blocks * (int(self.cryst_per_block) + 1):
            raise ValueError("det_w_pseudo must be det_per_ring (no pseudo detectors) or "
                             "blocks_per_ring * transaxial_multip * (cryst_per_block + 1) (one after every block)")
        if self.cr_p <= 0 or self.diameter <= 0 or self.rings <= 0 or self.cr_pz <= 0:
            raise ValueError("The scanner model needs cr_p, diameter, rings and cr_pz")
        if int(self.nLayers) > 1:
            raise ValueError("Multiple crystal layers (nLayers > 1) are not supported by this implementation")
        detw = int(self.det_w_pseudo)
        if self.Nang == 0:
            self.Nang = detw // 2
        if self.Ndist == 0:
            self.Ndist = detw // 2
        if int(self.Nang) > detw // 2:
            raise ValueError(f"Number of sinogram angles can be at most the number of detectors per ring divided "
                             f"by two ({detw // 2})!")
        if self.ring_difference <= 0:
            self.ring_difference = self.rings - 1
        xc, yc = G.omega_crystal_ring(self.blocks_per_ring, self.cryst_per_block, self.cr_p, self.diameter,
                                      self.transaxial_multip, self.DOI, self.flip_image, self.offangle, pseudo)
        self.x = G.omega_sinogram_xy(xc, yc, int(self.Ndist), int(self.Nang), int(self.ndist_side)).ravel()
        nz = int(np.asarray(self.Nz).ravel()[0])
        cpba = int(self.cryst_per_block_axial) or int(self.rings)
        z = G.omega_sinogram_z(int(self.rings), float(self.cr_pz), int(self.span), int(self.ring_difference), nz,
                               ring_gaps=self.ringGaps, cryst_per_block_axial=cpba)
        if self.NSinos > 0:
            z = z[: int(self.NSinos)]
        self.z = z.astype(np.float32).ravel()

    def addProjector(self):
        """Validate the configuration, place the grid and form the subsets (proj.py:520-1066)."""
        self.OMEGAErrorCheck()
        self._build_tables()
        Nx, Ny, Nz = (int(np.asarray(v).ravel()[0]) for v in (self.Nx, self.Ny, self.Nz))
        self.grid = G.Grid.from_fov(Nx, Ny, Nz, float(self.FOVa_x), float(self.FOVa_y), float(self.axial_fov),
                                    self.oOffsetX, self.oOffsetY, self.oOffsetZ)
        g = self.grid
        self.dx = np.array([g.dx], dtype=np.float32)
        self.dy = np.array([g.dy], dtype=np.float32)
        self.dz = np.array([g.dz], dtype=np.float32)
        pt = self.projector_type % 10 if self.projector_type > 10 else self.projector_type
        if pt == 3 and getattr(self, "V", None) is None:
            # computeVoxelVolumes (coordinates.py:190-205): the table and its scalars; voxel_radius becomes the
            # voxel sphere's radius in mm, as there
            self.V, self.bmin, self.bmax, self.Vmax, self.voxel_radius = G.voxel_volume_table(
                float(self.tube_radius), float(self.voxel_radius), self.dx[0].item(), self.dy[0].item(),
                self.dz[0].item())
        self.bx = np.array([g.bx], dtype=np.float32)
        self.by = np.array([g.by], dtype=np.float32)
        self.bz = np.array([g.bz], dtype=np.float32)
        self.Nx = np.array(Nx, dtype=np.uint32, ndmin=1)
        self.Ny = np.array(Ny, dtype=np.uint32, ndmin=1)
        self.Nz = np.array(Nz, dtype=np.uint32, ndmin=1)
        self.N = self.Nx.astype(np.uint64) * self.Ny.astype(np.uint64) * self.Nz.astype(np.uint64)
        if self.listmode:
            nlor = self.x.size // 6
            per_subset = np.full(self.subsets, nlor // self.subsets, dtype=np.int64)
            per_subset[-1] = nlor - per_subset[:-1].sum()
            self.index = np.arange(nlor, dtype=np.int64)
            kerroin = 1
        elif self.subsets > 1 and self.subsetType < 8:
            # measurement-level subsets: the library looks each measurement's bin up in the index table
            self.index, per_subset = G.subset_index_measurements(self.subsetType, self.subsets, int(self.nRowsD),
                                                                 int(self.nColsD), int(self.NSinos), self.seed)
            kerroin = 1
        else:
            st = self.subsetType if self.subsetType in (8, 9, 11) else 8
            self.index, per_subset = G.subset_index_projections(st, int(self.nProjections),
                                                                self.subsets if self.subsets > 1 else 1, int(self.seed))
            if self.subsets > 1:
                if self.CT:
                    self.x = np.ascontiguousarray(self.x.reshape(-1, 6)[self.index].reshape(-1))
                    w = 6 if self.pitch else 2
                    self.z = np.ascontiguousarray(self.z.reshape(-1, w)[self.index].reshape(-1))
                else:
                    self.z = np.ascontiguousarray(self.z.reshape(-1, 2)[self.index].reshape(-1))
            kerroin = self.nRowsD * self.nColsD
        self.nMeas = np.insert(np.cumsum(per_subset), 0, 0).astype(np.int64)
        self.nMeasSubset = (np.diff(self.nMeas) * kerroin).reshape(-1, 1).astype(np.int64)
        self.nProjSubset = np.diff(self.nMeas).reshape(-1, 1).astype(np.int64)
        self.nTotMeas = self.nMeas * kerroin
        if self.x0.size == 0:
            self.x0 = np.full(int(self.N[0]), 1e-4 if self.CT else 1.0, dtype=np.float32)
        self.x0 = np.asarray(self.x0, dtype=np.float32).ravel(order="F")
        if self.subsets > 1:
            self.subset = 0
        self.projectorAdded = True

    def _measurement_subsets(self):
        return self.subsets > 1 and not self.listmode and not self.CT and self.subsetType < 8

    def subset_order(self, data):
        """Reorder full measurement data (sinogram order) into subset order."""
        if self._measurement_subsets():
            return np.ascontiguousarray(np.asarray(data, dtype=np.float32).ravel()[self.index])
        per = self.nRowsD * self.nColsD
        d = np.asarray(data, dtype=np.float32).reshape(-1, per)
        return np.ascontiguousarray(d[self.index].reshape(-1))

    def _attenuation(self):
        """The attenuation table handed to the library and its mode (1 image, 2 per measurement).

        An image (CTAttenuation and CT_attenuation, the defaults) gets the reference's preparation
        (prepass.py:126-150): N values, or an (Nx, Ny, Nz) array; rotated by rotateAttImage quarter turns
        in the transaxial plane, flipped (flipAttImageXY, flipAttImageZ), divided by 10 when attIncm, and
        passed x fastest. The library computes each LOR's factor exp(-sum len * mu) once. Per-measurement
        factors (init.py:518-521, prepass.py:465-470) are reordered into subset order as the
        normalisation is."""
        v = np.asarray(self.vaimennus, dtype=np.float32)
        if v.size == 0:
            raise ValueError("attenuation_correction needs vaimennus: an attenuation image or one factor per measurement")
        g = self.grid
        if self.CTAttenuation and self.CT_attenuation:
            if v.size != g.N:
                raise ValueError(f"Attenuation data is of different size than the reconstructed image "
                                 f"({v.size} values, {g.N} voxels)")
            a = v.reshape((g.Nx, g.Ny, g.Nz), order="F") if v.ndim != 3 else v
            if self.rotateAttImage != 0:
                a = np.rot90(a, int(self.rotateAttImage))
            if self.flipAttImageXY:
                a = np.fliplr(a)
            if self.flipAttImageZ:
                a = np.flip(a, 2)
            if self.attIncm:
                a = a / np.float32(10.0)
            return np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel(order="F")), 1
        if self.listmode:
            return np.ascontiguousarray(v.ravel()), 2
        return self.subset_order(v), 2

    @property
    def nBins(self) -> int:
        """TOF bins of every measurement (1 without TOF)."""
        return int(self.TOF_bins) if self._tof() else 1

    def _tof(self) -> bool:
        """TOF is used when TOF_bins > 1 and TOF_bins_used equals it (0 meaning all bins); other
        TOF_bins_used values ask the reference to sum the bins, i.e. a non-TOF reconstruction
        (proj.py:681-691, :1208-1210). The reference accepts it for projector types 1-3 (proj.py:1247-1248);
        here initProj refuses it for any type but 1."""
        used = int(self.TOF_bins_used) or int(self.TOF_bins)
        return int(self.TOF_bins) > 1 and used == int(self.TOF_bins) and not self.CT

    def _tof_tables(self):
        """sigma_x (mm) and the bin centres (mm) the kernels get (proj.py:681-705, init errors :1247-1254)."""
        bins = int(self.TOF_bins)
        if self.TOFCenter.size == 0:
            if self.TOF_width <= 0:
                raise ValueError("TOF width (self.TOF_width) must be greater than zero.")
            if self.TOF_FWHM <= 0:
                raise ValueError("TOF FWHM (self.TOF_FWHM) must be greater than zero.")
            self.sigma_x = G.tof_sigma(self.TOF_FWHM)
            self.TOFCenter = G.tof_centers(bins, self.TOF_width, self.TOF_offset)
        elif self.sigma_x <= 0:
            if self.TOF_FWHM <= 0:
                raise ValueError("TOF FWHM (self.TOF_FWHM) must be greater than zero.")
            self.sigma_x = G.tof_sigma(self.TOF_FWHM)
        c = np.ascontiguousarray(self.TOFCenter, dtype=np.float32).ravel()
        if c.size != bins:
            raise ValueError(f"TOFCenter holds {c.size} values for {bins} TOF bins")
        return float(self.sigma_x), c

    def _descriptor(self):
        d = _lib.GeomDesc()
        d.struct_size = C.sizeof(_lib.GeomDesc)
        pt = self.projector_type % 10 if self.projector_type > 10 else self.projector_type
        d.projector_type = pt
        d.mode = _lib.MODE_CT_FLAT if self.CT else (_lib.MODE_LIST6 if self.listmode else _lib.MODE_PET_SINO)
        g = self.grid
        d.Nx, d.Ny, d.Nz = g.Nx, g.Ny, g.Nz
        d.dx, d.dy, d.dz = g.dx, g.dy, g.dz
        d.bx, d.by, d.bz = g.bx, g.by, g.bz
        d.size_x, d.size_y = int(self.nRowsD), int(self.nColsD)
        d.n_proj = int(self.nProjections)
        d.ct_pitch = int(bool(self.pitch))
        d.dPitchZ = float(self.dPitchY)
        d.global_factor = float(self.global_correction_factor)
        d.x, d.x_len = self.x.ctypes.data, self.x.size
        d.z, d.z_len = self.z.ctypes.data, self.z.size
        keep = [self.x, self.z]
        if self.normalization_correction and self.normalization.size > 0:
            nrm = self.subset_order(self.normalization) if not self.listmode else \
                np.ascontiguousarray(self.normalization, dtype=np.float32).ravel()
            keep.append(nrm)
            d.norm, d.norm_len = nrm.ctypes.data, nrm.size
        nr = (max(int(self.n_rays_transaxial), 1), max(int(self.n_rays_axial), 1))
        if nr != (1, 1) and not self.CT:
            if pt not in (1, 4):   # -DN_RAYS for FPType / BPType 1 and 4 (init.py:357-360)
                raise ValueError("Several rays per LOR are supported with improved Siddon (projector_type = 1) and "
                                 "the interpolation-based projector (projector_type = 4)")
            d.rays_xy, d.rays_z = nr
            d.cr_xy, d.cr_z = float(self.cr_pz), float(self.cr_p)
        if self._tof() and pt not in (1, 3, 4):
            # proj.py:1247-1248
            raise ValueError("TOF is currently only supported with improved Siddon (projector_type = 1), "
                             "interpolation-based projector (projector_type = 4) and volume of intersection "
                             "(projector_type = 3)")
        if self._tof():
            sig, tc = self._tof_tables()
            keep.append(tc)
            d.tof_bins, d.tof_sigma, d.tof_epps = int(self.TOF_bins), sig, float(self.epps)
            d.tof_centers = tc.ctypes.data
        if self.additionalCorrection and not self.CT:
            if self.corrVector.size == 0:
                raise ValueError("additionalCorrection needs corrVector (one factor per measurement)")
            cv = self.subset_order(self.corrVector) if not self.listmode else \
                np.ascontiguousarray(self.corrVector, dtype=np.float32).ravel()
            keep.append(cv)
            d.scat, d.scat_len = cv.ctypes.data, cv.size
        if self.attenuation_correction and not self.CT:
            atn, d.atten_mode = self._attenuation()
            keep.append(atn)
            d.atten, d.atten_len = atn.ctypes.data, atn.size
        if pt == 2:
            # the kernel's tube width is tube_width_z (init.py:652); the sweeps follow orthTransaxial /
            # orthAxial (proj.py:1041-1054 -> -DCRYSTXY / -DCRYSTZ, init.py:286-290)
            if not self.tube_width_z > 0:
                raise ValueError("The orthogonal projector needs tube_width_z > 0: it is the kernel's tube width "
                                 "(with tube_width_z = 0 every voxel is outside the tube)")
            d.orth_width = float(self.tube_width_z)
            d.orth_flags = int(self.tube_width_xy > 0) | (int(self.tube_width_z > 0) << 1)
        if pt == 3:
            # the kernel's radius is tube_radius (init.py:653), both sweeps on (proj.py:1041-1054); bmin,
            # bmax and Vmax reach the kernel as float32 (init.py:654)
            if not self.tube_radius > 0:
                raise ValueError("The volume-of-intersection projector needs tube_radius > 0")
            V = np.ascontiguousarray(self.V, dtype=np.float32)
            keep.append(V)
            d.orth_width = float(np.float32(self.tube_radius))
            d.orth_flags = 3
            d.vol_V, d.vol_len = V.ctypes.data, V.size
            d.vol_bmin, d.vol_bmax, d.vol_vmax = float(self.bmin), float(self.bmax), float(self.Vmax)
        if getattr(self, "voxel_backprojection", False):
            d.reserved[2] = 1
        if self._measurement_subsets():
            idx = np.ascontiguousarray(self.index, dtype=np.uint32)
            keep.append(idx)
            d.meas_index, d.meas_index_len = idx.ctypes.data, idx.size
        d.dPitchX = float(self.dPitchX)
        d.dPitchY = float(self.dPitchY)
        if pt == 4:
            d.interp_dl, d.interp_kerroin = self._interp_scalars()
        return d, keep

    def _interp_scalars(self):
        """interp_dl (mm) and interp_kerroin of projector type 4 (computeProjectorScalingValues,
        coordinates.py:241-262): dL = FOVa_x / Nx, or dL * FOVa_x / Nx when dL is set (a negative dL, the
        Joseph variant, stays negative); kerroin = (dx dy dz) / (dPitchX dPitchY sourceToDetector) for CT,
        sourceToDetector taken from the first projection's source-panel distance when it is not set."""
        vox = float(self.FOVa_x) / float(np.asarray(self.Nx).ravel()[0])
        dl = vox if self.dL == 0 else float(self.dL) * vox
        kerroin = 0.0
        if self.CT:
            sdd = float(self.sourceToDetector)
            if sdd <= 0:
                t = np.asarray(self.x, dtype=np.float64).reshape(-1, 6)[0]
                sdd = float(np.linalg.norm(t[3:] - t[:3]))
            if self.dPitchX <= 0 or self.dPitchY <= 0 or sdd <= 0:
                raise ValueError("The interpolation projector on CT data needs dPitchX, dPitchY and the source-detector "
                                 "distance")
            g = self.grid
            kerroin = (float(g.dx) * float(g.dy) * float(g.dz)) / (float(self.dPitchX) * float(self.dPitchY) * sdd)
        return float(np.float32(dl)), float(np.float32(kerroin))

    def initProj(self):
        """Create the device-side geometry (proj.py:1913 -> init.py:39)."""
        if not self.projectorAdded:
            self.addProjector()
        d, keep = self._descriptor()
        self._geom = Geometry(d, keep)
        self.projectorInitialized = True

    # ------------------------------------------------------------------ operators
    def _range(self, subset):
        if subset == -1:
            subset = self.subset
        if self.subsets == 1:
            subset = 0
        first = int(self.nMeas[subset])
        nproj = int(self.nMeas[subset + 1] - self.nMeas[subset])
        per = 1 if (self.listmode or self._measurement_subsets()) else self.nRowsD * self.nColsD
        return first, nproj * per

    def forwardProject(self, f, subset=-1):
        """y = A_subset f (projfunctions.forwardProjection, projfunctions.py:86)."""
        import torch
        if not self.projectorInitialized:
            self.initProj()
        if isinstance(f, list):
            raise ValueError("Multi-resolution volumes are not supported by this implementation!")
        first, n = self._range(subset)
        f = self._as_image(f)
        y = torch.empty(n * self.nBins, dtype=torch.float32, device=f.device)
        _lib.call("omg_forward_project", self._geom.handle, first, n, C.c_void_p(f.data_ptr()),
                  C.c_void_p(y.data_ptr()), _stream_ptr(torch))
        return y

    def backwardProject(self, y, subset=-1, sens=None):
        """f = A_subset^T y (projfunctions.backwardProjection, projfunctions.py:735).

        ``sens`` (optional, flat float32 image tensor) accumulates A^T 1 in the same pass.
        """
        import torch
        if not self.projectorInitialized:
            self.initProj()
        first, n = self._range(subset)
        y = self._as_meas(y, n * self.nBins)
        f = torch.zeros(int(self.N[0]), dtype=torch.float32, device=y.device)
        sp = C.c_void_p(0)
        if sens is not None:
            if sens.dtype != torch.float32 or not sens.is_contiguous() or sens.numel() != int(self.N[0]):
                raise ValueError("sens must be a contiguous float32 tensor with N elements")
            sp = C.c_void_p(sens.data_ptr())
        _lib.call("omg_backward_project", self._geom.handle, first, n, C.c_void_p(y.data_ptr()),
                  C.c_void_p(f.data_ptr()), sp, _stream_ptr(torch))
        return f

    def computeSensitivityImage(self, crystal_xy, ring_z, max_lors: int = 1 << 25):
        """Sensitivity image of list-mode data (compute_sensitivity_image, proj.py:281; the reference's
        -DSENS backprojection, ProjectorClass.h:3217-3264,3913-3978): the backprojection of ones over
        every ordered crystal pair of every ordered ring pair, with this projector's type, tube settings,
        rays, global correction factor and attenuation image (per-event tables -- normalisation,
        multiplicative corrections, attenuation factors -- do not describe these LORs).
        crystal_xy: (det_per_ring, 2) crystal centres; ring_z: (rings,) ring positions. The LORs run
        as a sinogram (geometry.listmode_sensitivity_tables), at most max_lors per launch."""
        import torch
        if not self.projectorInitialized:
            self.initProj()
        if self.CT:
            raise ValueError("The sensitivity image of list-mode data is for PET")
        if self._tof():
            raise ValueError("The list-mode sensitivity image is computed without TOF")
        x, z = G.listmode_sensitivity_tables(crystal_xy, ring_z)
        det = x.size // 4
        det = int(round(math.sqrt(det)))
        planes = z.size // 2
        d, keep = self._descriptor()
        d.mode = _lib.MODE_PET_SINO
        d.size_x = d.size_y = det
        d.n_proj = planes
        d.x, d.x_len = x.ctypes.data, x.size
        d.z, d.z_len = z.ctypes.data, z.size
        d.norm, d.norm_len = None, 0
        d.meas_index, d.meas_index_len = None, 0
        if d.atten_mode == 2:
            raise ValueError("The list-mode sensitivity image takes attenuation from an attenuation image "
                             "(CTAttenuation); per-event attenuation factors do not describe its LORs")
        # an attenuation image stays: the library computes every sensitivity LOR's factor from it, as the
        # reference's -DSENS kernel integrates it along each LOR (projectorType123.cl, ProjectorClass.h:3937-3953)
        d.scat, d.scat_len = None, 0
        geom = Geometry(d, keep + [x, z])
        dev = torch.device("cuda")
        sens = torch.zeros(int(self.N[0]), dtype=torch.float32, device=dev)
        per = det * det
        chunk = max(1, min(planes, max_lors // per))
        ones = torch.ones(per * chunk, dtype=torch.float32, device=dev)
        try:
            for p0 in range(0, planes, chunk):
                n = per * min(chunk, planes - p0)
                # sinogram calls start at a plane: `first` is p0
                _lib.call("omg_backward_project", geom.handle, p0, n, C.c_void_p(ones.data_ptr()),
                          C.c_void_p(sens.data_ptr()), C.c_void_p(0), _stream_ptr(torch))
        finally:
            torch.cuda.synchronize()
            geom.close()
        return sens

    def T(self):
        self.trans = True
        return self

    def __mul__(self, B):
        if self.trans:
            self.trans = False
            return self.backwardProject(B, self.subset)
        return self.forwardProject(B, self.subset)

    def __rmul__(self, B):
        return self

    # ------------------------------------------------------------------ helpers
    def _as_image(self, f):
        import torch
        if not isinstance(f, torch.Tensor):
            f = torch.as_tensor(np.asarray(f, dtype=np.float32).ravel(order="F"), device="cuda")
        if f.dtype != torch.float32:
            raise ValueError("images must be float32")
        if not f.is_cuda:
            raise ValueError("images must be device tensors (useTorch=True, useCUDA=True)")
        f = f.reshape(-1).contiguous()
        if f.numel() != int(self.N[0]):
            raise ValueError(f"image has {f.numel()} elements, expected {int(self.N[0])}")
        return f

    def _as_meas(self, y, n):
        import torch
        if not isinstance(y, torch.Tensor):
            y = torch.as_tensor(np.asarray(y, dtype=np.float32).ravel(), device="cuda")
        if y.dtype != torch.float32:
            raise ValueError("measurements must be float32")
        if not y.is_cuda:
            raise ValueError("measurements must be device tensors (useTorch=True, useCUDA=True)")
        y = y.reshape(-1).contiguous()
        if y.numel() != n:
            raise ValueError(f"measurement vector has {y.numel()} elements, expected {n} for this subset")
        return y

    @property
    def geometry(self) -> Geometry:
        if not self.projectorInitialized:
            self.initProj()
        return self._geom
