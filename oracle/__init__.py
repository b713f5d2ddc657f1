"""CPU oracle of the CSD / ICLS path — test infrastructure only (never imported by mrtrix3_amd)."""
