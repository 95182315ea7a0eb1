"""Problem scripts of the benchmark configurations, written against
opensbli_amd.api (OpenSBLI's problem-definition names)."""
