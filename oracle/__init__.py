"""CPU oracle of the splat frame pipeline (test infrastructure only)."""
