"""CPU oracle of the reference algorithms: test infrastructure, never imported by the product."""
