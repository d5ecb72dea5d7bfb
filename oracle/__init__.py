"""CPU oracle of the PBSM3D step: test infrastructure, not part of the product."""
