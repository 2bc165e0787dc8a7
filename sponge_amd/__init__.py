"""MI355X-native SPONGE nonbonded hot path (neighbor list, LJ + PME direct
pair kernel with exclusions, PME reciprocal)."""
