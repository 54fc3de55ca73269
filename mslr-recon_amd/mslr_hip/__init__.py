"""mslr_hip: MI355X (gfx950) MSLR reconstruction hot path behind the reference's API.

Modules mirror the reference's Julia modules:
    sense       Asense_hip / SenseOp / block_diag     (src/sense_gpu.jl, scripts/reconstruct.jl:133-142)
    recon       img2patches, patches2img, SVST, patchSVST, patch_nucnorm   (src/recon.jl)
    mirt_mod    POGM device loop, poweriter_mod       (src/mirt_mod.jl)
    reconstruct run_recon / run_recon_arrays          (scripts/reconstruct.jl)
    synth       deterministic synthetic problems
All compute runs in libmslr_hip.so (HIP, gfx950); there is no CPU fallback.
"""
__all__ = ["sense", "recon", "mirt_mod", "reconstruct", "synth", "device"]
