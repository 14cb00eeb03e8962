"""CPU oracle of the RCA delay-and-sum beamformer (test infrastructure only)."""
