"""MI355X-native SHCI variational step (drop-in for the variational path of
sanshar/Dice): HIP kernels behind the C ABI in include/shci_gpu.h, with the
Python mirror of the reference interface in dice_amd.shci."""
__all__ = ["shci", "synth"]
