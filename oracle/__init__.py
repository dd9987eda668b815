"""CPU oracle of the RoadRunner path (test infrastructure only; never imported by pytransit_amd)."""
